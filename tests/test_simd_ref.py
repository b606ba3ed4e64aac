"""The CPU baseline's GFNI / AVX-512 restatement (oracle/simd_ref.c) against the
scalar oracle, byte for byte, before bench.py times it.  Skipped on CPUs
without GFNI and AVX-512."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "oracle"))
sys.path.insert(0, os.path.join(HERE, "..", "reed-solomon_amd"))
import oracle as O  # noqa: E402
import fftgen  # noqa: E402

pytestmark = pytest.mark.skipif(not O.simd_available(), reason="CPU lacks GFNI / AVX-512")

LENGTHS = [1, 63, 64, 65, 127, 128, 129, 255, 256, 257, 1000, 4096 + 17]


@pytest.mark.parametrize("nout,k", [(1, 1), (4, 10), (6, 7), (7, 13), (13, 20), (16, 33)])
def test_apply_matches_scalar(nout, k):
    rng = np.random.default_rng(nout * 100 + k)
    rows = rng.integers(0, 256, (nout, k), dtype=np.uint8)
    for length in LENGTHS:
        ins = [rng.integers(0, 256, length, dtype=np.uint8) for _ in range(k)]
        want = O.rs8_apply(rows, ins, length)
        got = O.simd_rs8_apply(rows, ins, length)
        assert (got == want).all(), (nout, k, length)


@pytest.mark.parametrize("k,m", sorted(fftgen.GENERATED_SHAPES))
def test_fft_programs_match_matrix_encode(k, m):
    p = fftgen.build_encode_program(k, m)
    prog = O.simd_fft_program(p.ops, p.parity_regs, p.register_count)
    G = O.build_matrix(8, k, k + m)
    for length in LENGTHS:
        data = O.fill(0x5EED + length, k * length).reshape(k, length)
        assert (O.simd_fft_encode(prog, k, data) == O.rs8_encode(G, k, data)).all(), length


@pytest.mark.parametrize("k,m", sorted(fftgen.GENERATED_SHAPES))
def test_generated_straight_line_encoders(k, m):
    G = O.build_matrix(8, k, k + m)
    for length in LENGTHS:
        data = O.fill(0xFEED + length, k * length).reshape(k, length)
        assert (O.simd_fft_generated(k, m, data) == O.rs8_encode(G, k, data)).all(), length
    assert O.simd_fft_generated(10, 4, np.zeros((10, 64), np.uint8)) is None


@pytest.mark.parametrize("k,m", [(10, 4), (7, 13), (32, 32), (3, 2)])
def test_reconstruct_round_trip(k, m):
    rng = np.random.default_rng(k + m)
    G = O.build_matrix(8, k, k + m)
    for length in (1, 129, 5000):
        data = O.fill(k * 1000 + length, k * length).reshape(k, length)
        full = np.concatenate([data, O.rs8_encode(G, k, data)])
        for data_only in (False, True):
            e = int(rng.integers(1, m + 1))
            erased = set(rng.choice(k + m, e, replace=False).tolist())
            cur = full.copy()
            for i in erased:
                cur[i] = 0
            present = [i not in erased for i in range(k + m)]
            assert O.simd_rs8_reconstruct(G, k, m, cur, present, data_only) == 0
            for i in range(k + m):
                if data_only and i >= k and i in erased:
                    assert not cur[i].any()
                else:
                    assert (cur[i] == full[i]).all(), (k, m, length, sorted(erased), i)
    # fewer than k present
    cur = np.zeros((k + m, 16), np.uint8)
    assert O.simd_rs8_reconstruct(G, k, m, cur, [i < k - 1 for i in range(k + m)]) == 10


@pytest.mark.parametrize("k,m", [(10, 10), (10, 4), (32, 32)])
def test_prepared_calls(k, m):
    """The bench's timed calls: encode (generated program or fused apply) and
    the two reconstruct applies from a cached plan."""
    G = O.build_matrix(8, k, k + m)
    length = 3000
    data = O.fill(77 + k, k * length).reshape(k, length)
    parity = np.zeros((m, length), np.uint8)
    O.simd_encode_call(G, k, m, data, parity)()
    assert (parity == O.rs8_encode(G, k, data)).all()
    full = np.concatenate([data, parity])
    erased = {0, k - 1, k, k + m - 1}
    cur = full.copy()
    for i in erased:
        cur[i] = 0
    plan = O.rs8_decode_plan(G, k, m, [i not in erased for i in range(k + m)])
    for call in O.simd_reconstruct_calls(G, k, m, cur, plan):
        call()
    assert (cur == full).all()
