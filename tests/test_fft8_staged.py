"""GF(2^8) staged FFT program (src/fft.rs:393-592) on the CPU.

* The library's C++ builder, lowered for the device interpreter, replays to
  the generator-matrix parity of the oracle on the reference's 21 program
  shapes (src/fft.rs:979-1005, test program_matches_matrix :1008-1041) and
  on the staged-route shapes, from a register file of stale values.  Every
  constant is a byte of the subfield (Karatsuba word [c, c, 0]), which is
  what lets the byte-plane kernel use the word's low byte alone: the replay
  runs the tower arithmetic on byte inputs and the high byte of every output
  must stay zero.
* The multiply and register counts of src/fft.rs's lowering are pinned by
  their closed form for k = m = 2^j: W_l vanishes exactly on [0, 2^l), so the
  base transforms at beta 0 skip one block per level and the chunk at beta k
  skips none, giving 3 (k/2) log2 k - 2 (k - 1) multiplies over 2k registers.
* The route rule is the reference's (src/reedsolomon.rs:204-234): power-of-two
  k >= 2, not a generated shape, at most 160 registers, 2 * mults < k * m.
* The GF(2^8) FFT erasure decoder (formal-derivative method over the byte
  field, no reference counterpart) replays to the oracle's codewords for
  random erasure patterns.
"""
import ctypes
import math

import numpy as np
import pytest

import oracle as O
from test_fft16 import _lowered_from, replay

# src/fft.rs:979-1005
REFERENCE_SHAPES = [(1, 1), (1, 4), (2, 2), (3, 5), (4, 2), (5, 3), (6, 4), (7, 13), (9, 3),
                    (10, 10), (10, 7), (11, 5), (13, 7), (14, 14), (16, 4), (17, 17), (18, 6),
                    (20, 10), (24, 8), (32, 32), (37, 21)]
# top_of_field_counts (src/fft.rs:1043-1070) and staged-route shapes
EXTRA_SHAPES = [(64, 64), (128, 128), (200, 55), (200, 20), (230, 25), (250, 5), (100, 156),
                (128, 32), (64, 16), (8, 8), (32, 8), (2, 254)]
GENERATED = [(7, 13), (10, 10), (14, 14), (16, 16), (18, 6), (32, 32)]


def _info(trs, k, m):
    return trs.fft8_program_info(k, m)


def _lowered8(trs, k, m):
    return _lowered_from(trs._lib.lib.trs_fft8_lowered_program, k, m)


def _byte_parity_by_replay(trs, k, m, data):
    """Runs the lowered byte program over (k, L) data bytes with the tower
    arithmetic; a byte constant's product never leaves the low byte."""
    steps, consts, ops, parity, regs = _lowered8(trs, k, m)
    out = replay(steps, consts, ops, np.arange(k), parity, regs, data.astype(np.int64))
    assert (out >> 8 == 0).all()
    return out.astype(np.uint8)


@pytest.mark.parametrize("k,m", REFERENCE_SHAPES + EXTRA_SHAPES)
def test_byte_program_replays_to_matrix(trs, k, m):
    rng = np.random.default_rng(k * 257 + m)
    data = rng.integers(0, 256, size=(k, 64), dtype=np.uint8)
    G = O.build_matrix(8, k, k + m)
    assert (_byte_parity_by_replay(trs, k, m, data) == O.rs8_encode(G, k, data)).all()


def test_constants_are_subfield_words(trs):
    # every multiply constant word is [c, c, 0] (+ flags): the low and the high
    # plane see the same byte product
    steps, consts, ops, _, _ = _lowered8(trs, 64, 64)
    words = consts & 0xFFFFFF
    assert ((words & 0xFF) == ((words >> 8) & 0xFF)).all()
    assert ((words >> 16) == 0).all()
    kinds = ops[1::2] & 0xFF
    mul = (kinds == 2) | (kinds == 3)
    cw = ops[1::2][mul] >> 8
    assert ((cw & 0xFF) == ((cw >> 8) & 0xFF)).all() and ((cw >> 16) == 0).all()


@pytest.mark.parametrize("j", [1, 2, 3, 4, 5, 6])
def test_square_power_of_two_counts_closed_form(trs, j):
    k = 1 << j
    info = _info(trs, k, k)
    assert info["mult_count"] == 3 * (k // 2) * j - 2 * (k - 1)
    assert info["register_count"] == 2 * k
    # the device lowering drops the base transform: no parity lands in it
    assert info["device_mult_count"] == 3 * (k // 2) * j - 2 * (k - 1) - ((k // 2) * j - (k - 1))


def _rule(info, k, m):
    return (k >= 2 and k & (k - 1) == 0 and (k, m) not in GENERATED
            and info["register_count"] <= 160 and 2 * info["mult_count"] < k * m)


@pytest.mark.parametrize("k,m", REFERENCE_SHAPES + EXTRA_SHAPES + GENERATED +
                         [(64, 63), (4, 4), (2, 1), (128, 33), (128, 64), (64, 192)])
def test_reference_route_rule(trs, k, m):
    info = _info(trs, k, m)
    assert info["reference_route"] == int(_rule(info, k, m))


def test_reference_route_examples(trs):
    assert _info(trs, 64, 64)["reference_route"] == 1
    assert _info(trs, 8, 8)["reference_route"] == 1
    assert _info(trs, 16, 16)["reference_route"] == 0   # generated shape
    assert _info(trs, 12, 8)["reference_route"] == 0    # k not a power of two
    assert _info(trs, 1, 4)["reference_route"] == 0     # k < 2
    assert _info(trs, 128, 128)["reference_route"] == 0  # 256 registers > 160
    assert _info(trs, 128, 32)["register_count"] == 160
    assert _info(trs, 0, 4) is None and _info(trs, 200, 57) is None


def test_program_info_group_bytes(trs):
    info = _info(trs, 64, 64)
    assert info["group_bytes"] == 32 * info["device_regs"]
    assert info["device_regs"] <= info["register_count"]
    assert math.isfinite(info["device_steps"]) and info["device_steps"] > 0


# ---- additive-FFT erasure decoder over GF(2^8) -----------------------------

def _decode_program8(trs, k, n, present):
    lib = trs._lib.lib
    flags = (ctypes.c_uint8 * n)(*[1 if p else 0 for p in present])
    counts = (ctypes.c_size_t * 6)()
    null = ctypes.POINTER(ctypes.c_uint32)()
    rc = lib.trs_fft8_decode_program(k, n, flags, null, 0, null, 0, null, 0, null, 0, null, 0,
                                     counts)
    assert rc == 0
    bufs = [np.zeros(max(1, int(c)), np.uint32) for c in counts[:5]]
    ptr = [b.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)) for b in bufs]
    rc = lib.trs_fft8_decode_program(k, n, flags, ptr[0], bufs[0].size, ptr[1], bufs[1].size,
                                     ptr[2], bufs[2].size, ptr[3], bufs[3].size, ptr[4],
                                     bufs[4].size, counts)
    assert rc == 0
    out = [b[:c].astype(np.int64) for b, c in zip(bufs, counts[:5])]
    return out[0].reshape(-1, 8), out[1], out[2], out[3], out[4], int(counts[5])


@pytest.mark.parametrize("k,m", [(2, 2), (4, 2), (3, 13), (7, 13), (10, 4), (10, 10), (17, 17),
                                 (32, 32), (64, 64), (100, 156), (200, 55), (128, 128), (255, 1),
                                 (1, 255)])
def test_byte_fft_decoder_replays_to_codeword(trs, k, m):
    """The GF(2^8) decoder program rebuilds every erased shard (data and
    parity) of oracle-encoded stripes, for random erasure patterns of 1..m
    shards; the tower arithmetic on byte inputs keeps every value a byte."""
    n = k + m
    rng = np.random.default_rng(k * 31 + m)
    data = rng.integers(0, 256, size=(k, 24), dtype=np.uint8)
    cw = np.concatenate([data, O.rs8_encode(O.build_matrix(8, k, n), k, data)]).astype(np.int64)
    for e in sorted({1, m, int(rng.integers(1, m + 1))}):
        erased = set(rng.choice(n, e, replace=False).tolist())
        present = [i not in erased for i in range(n)]
        steps, consts, ops, load, store, regs = _decode_program8(trs, k, n, present)
        assert list(load) == [i for i in range(n) if present[i]][:k]
        assert list(store) == sorted(i for i in erased if i < k) + sorted(i for i in erased if i >= k)
        got = replay(steps, consts, ops, load, store, regs, cw[load])
        assert (got >> 8 == 0).all()
        assert (got == cw[store]).all(), (k, m, sorted(erased))
