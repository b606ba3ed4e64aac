"""CPU pins for the FFT program builder (reed-solomon_amd/fftgen.py).

* The six generated programs equal the reference's src/fft_programs.rs op
  lists byte for byte (SHA-256 of the token lines, golden fixture) along with
  their register / op / multiply counts (src/fft_programs.rs:19-787 headers).
* Replaying any program over random byte columns gives the generator-matrix
  parity of the CPU oracle (src/fft.rs:984-1041 does the same comparison).
* WHAT tables: zero on each level's span, 1 at the pivot (src/fft.rs:890-909).
"""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "reed-solomon_amd"))
import fftgen  # noqa: E402

KATS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))


@pytest.mark.parametrize("shape", KATS["fft_generated_programs"]["shapes"],
                         ids=lambda s: f"{s['k']}_{s['m']}")
def test_generated_program_matches_reference(shape):
    p = fftgen.build_encode_program(shape["k"], shape["m"])
    assert p.register_count == shape["registers"]
    assert len(p.ops) == shape["ops"]
    assert p.mult_count == shape["multiplies"]
    lines = p.text_lines()
    assert len(lines) == shape["op_lines"]
    assert hashlib.sha256("\n".join(lines).encode()).hexdigest() == shape["op_sha256"]


def test_multiply_counts_from_bench_results():
    # BENCH-RESULTS.md:26 -- 29 vs 91 (7,13) and 34 vs 100 (10,10) multiplies
    assert fftgen.build_encode_program(7, 13).mult_count == 29
    assert fftgen.build_encode_program(10, 10).mult_count == 34


def test_what_table_shape():
    for level in range(fftgen.LEVELS):
        assert all(fftgen.WHAT[level][p] == 0 for p in range(1 << level))
        assert fftgen.WHAT[level][1 << level] == 1


@pytest.mark.parametrize("k,m", [(7, 13), (10, 10), (14, 14), (16, 16), (18, 6), (32, 32),
                                 (4, 2), (10, 4), (12, 8), (24, 8), (5, 27), (64, 64), (3, 100)])
def test_program_parity_equals_matrix_parity(k, m):
    p = fftgen.build_encode_program(k, m)
    G = O.build_matrix(8, k, k + m)
    rng = np.random.default_rng(k * 1000 + m)
    data = rng.integers(0, 256, (k, 64)).astype(np.uint8)
    want = O.rs8_encode(G, k, data)
    for col in range(64):
        got = fftgen.run_program_bytes(p, [int(x) for x in data[:, col]])
        assert got == [int(x) for x in want[:, col]], (k, m, col)
