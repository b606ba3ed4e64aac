"""Tower-FFT route (src/fft16.rs) on the CPU.

* The oracle's restated builder replays to the matrix parity on the
  reference's staged_matches_matrix shapes with stale registers
  (src/fft16.rs:958-995), and keeps the (86,170) bound (:1077-1083).
* The library's C++ builder gives the same program statistics as the oracle
  (register, multiply, glue, stage and zero-list counts, routing decision).
* The device lowering the GPU interpreter runs (glue scheduled into parallel
  sub-steps, whole-range copies coalesced, Karatsuba constant words) replays
  on the CPU to the same parity, with every sub-step checked for conflicts.
"""
import ctypes

import numpy as np
import pytest

import fft16_ref as F
import oracle as O

# src/fft16.rs:966-979
REFERENCE_SHAPES = [(2, 2), (3, 5), (5, 3), (7, 13), (16, 16), (17, 33), (32, 96), (48, 144),
                    (63, 65), (86, 170), (100, 50), (128, 128)]
EXTRA_SHAPES = [(4, 2), (10, 4), (11, 20), (10, 10), (33, 31), (200, 56), (255, 1), (2, 1),
                (64, 192), (96, 32), (300, 212), (513, 511), (1000, 24)]


def _symbols(seed, k, n):
    rng = np.random.default_rng(seed)
    return rng.integers(0, 65536, size=(k, n), dtype=np.int64)


def _matrix_parity(k, m, sym):
    """Symbol-wise matrix oracle (src/fft16.rs:941-955): parity rows of the
    systematic generator applied to u16 symbol columns."""
    G = O.build_matrix(16, k, k + m)
    n = sym.shape[1]
    data = np.concatenate([sym & 0xFF, sym >> 8], axis=1).astype(np.uint8)  # split planes
    par = O.rs16_symbolwise(G[k:], list(data), 2 * n)
    return par[:, :n].astype(np.int64) | (par[:, n:].astype(np.int64) << 8)


@pytest.mark.parametrize("k,m", REFERENCE_SHAPES)
def test_oracle_program_matches_matrix(k, m):
    p = F.build_staged_program16(k, m)
    sym = _symbols(k * 1000 + m, k, 8)
    assert (F.run_symbols(p, sym).astype(np.int64) == _matrix_parity(k, m, sym)).all()


def test_oracle_window_bound():
    # src/fft16.rs:1077-1083
    p = F.build_staged_program16(86, 170)
    assert p.mult_count < 1100
    assert p.register_count <= F.MAX_STAGED_REGISTERS16


def _cpp_info(trs, k, m):
    return trs.fft16_program_info(k, m)


@pytest.mark.parametrize("k,m", REFERENCE_SHAPES + EXTRA_SHAPES + [(1024, 1024)])
def test_cpp_builder_matches_oracle(trs, k, m):
    p = F.build_staged_program16(k, m)
    info = _cpp_info(trs, k, m)
    assert info is not None
    assert info["register_count"] == p.register_count
    assert info["mult_count"] == p.mult_count
    assert info["glue"] == len(p.glue)
    assert info["stages"] == len(p.stages)
    assert info["zero_regs"] == len(p.zero_regs)
    assert info["profitable"] == int(F.profitable(p, k, m))
    assert info["device_regs"] <= info["register_count"]


def test_routing_rule_matches_reference_crossover(trs):
    # src/fft16.rs:78-80: the measured crossover of the multiply-count rule is
    # (11,20) on the matrix and (17,33) on the FFT.
    assert _cpp_info(trs, 11, 20)["profitable"] == 0
    assert _cpp_info(trs, 17, 33)["profitable"] == 1
    assert _cpp_info(trs, 1024, 1024)["profitable"] == 1
    # outside 2..=1024 data shards there is no program (src/reedsolomon16.rs:302)
    assert _cpp_info(trs, 1, 5) is None
    assert _cpp_info(trs, 1025, 5) is None


def test_c3_program_coalesces_to_1024_registers(trs):
    # interpolate = one inverse 1024-point transform, the parity chunk at 1024
    # copies the base (full-size fold) and transforms forward: after merging
    # the copies into the base registers the device file is 1024 registers,
    # 64 KiB of LDS per 32-column group.
    info = _cpp_info(trs, 1024, 1024)
    assert info["register_count"] == 2048
    assert info["device_regs"] == 1024
    assert info["group_bytes"] == 65536


# ---- replay of the device lowering ----------------------------------------

SKIP, ZERO, ONE = 1 << 24, 2 << 24, 4 << 24
K_XOR, K_COPY, K_MULXOR, K_MULCOPY, K_ZERO, K_DERIV = range(6)


def _lowered_from(fn, k, m):
    """(steps, consts, ops, parity registers, register count) of a lowered
    program dump (trs_fft16_lowered_program or trs_fft8_lowered_program)."""
    counts = (ctypes.c_size_t * 5)()
    null = ctypes.POINTER(ctypes.c_uint32)()
    assert fn(k, m, null, 0, null, 0, null, 0, null, 0, counts) == 0
    bufs = [np.zeros(max(1, int(c)), np.uint32) for c in counts[:4]]
    ptr = [b.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)) for b in bufs]
    rc = fn(k, m, ptr[0], bufs[0].size, ptr[1], bufs[1].size, ptr[2], bufs[2].size, ptr[3],
            bufs[3].size, counts)
    assert rc == 0
    steps = bufs[0][:counts[0]].reshape(-1, 8).astype(np.int64)
    return (steps, bufs[1][:counts[1]].astype(np.int64), bufs[2][:counts[2]].astype(np.int64),
            bufs[3][:counts[3]].astype(np.int64), int(counts[4]))


def _lowered(trs, k, m):
    return _lowered_from(trs._lib.lib.trs_fft16_lowered_program, k, m)


def _kara(cw, x):
    """The three-product tower multiply on Karatsuba bytes, as the kernel runs
    it: lo = c_lo x_lo ^ (c_hi L) x_hi, hi = c_lo x_lo ^ (c_hi ^ c_lo)(x_hi ^ x_lo)."""
    a0, a1, a2 = cw & 0xFF, (cw >> 8) & 0xFF, (cw >> 16) & 0xFF
    xl, xh = x & 0xFF, x >> 8
    t0 = F.gf8_mul(a0, xl)
    return ((t0 ^ F.gf8_mul(a1, xh ^ xl)) << 8) | (t0 ^ F.gf8_mul(a2, xh))


def replay(steps, consts, ops, load, store, regs, inputs):
    """Runs a device step list the way the kernel does (each step's items in
    parallel) from a register file of stale values: inputs[i] (symbol
    columns) goes to register load[i]; returns the registers of `store`."""
    n = inputs.shape[1]
    R = np.full((regs, n), 0x5A5A, np.int64)
    R[load] = inputs
    for kind, count, a, b, c, d, _, _ in steps:
        o = np.arange(count)
        if kind == 0:
            block = o >> b
            cw = consts[c + block][:, None]
            lo = a + (block << (b + 1)) + (o & ((1 << b) - 1))
            hi = lo + (1 << b)
            live = (cw[:, 0] & SKIP) == 0
            lo, hi, cw = lo[live], hi[live], cw[live]
            assert len(set(lo) | set(hi)) == 2 * len(lo)
            L, H = R[lo], R[hi]
            if d:
                H = H ^ L
            prod = np.where(cw & ONE, H, np.where(cw & ZERO, 0, _kara(cw & 0xFFFFFF, H)))
            L = L ^ prod
            if not d:
                H = H ^ L
            R[lo], R[hi] = L, H
        else:
            w0, w1 = ops[2 * (a + o)], ops[2 * (a + o) + 1]
            dst, src, knd, cw = w0 & 0xFFFF, w0 >> 16, w1 & 0xFF, (w1 >> 8)[:, None]
            # independent: distinct destinations, no destination read by another op
            assert len(set(dst)) == len(dst)
            reads_other = (knd != K_ZERO) & (knd != K_DERIV) & (src != dst)
            assert not (set(dst) & set(src[reads_other]))
            deriv = knd == K_DERIV
            if deriv.any():
                # dst ^= registers dst | 2^j for the clear bits j below the level count
                assert deriv.all()
                srcs = [[t | (1 << j) for j in range(int(c)) if not (t >> j) & 1]
                        for t, c in zip(dst, cw[:, 0])]
                assert not (set(dst) & {x for xs in srcs for x in xs})
                val = np.zeros((len(dst), R.shape[1]), np.int64)
                for i, xs in enumerate(srcs):
                    for x in xs:
                        val[i] ^= R[x]
                R[dst] ^= val
                continue
            S = R[src]
            val = np.where((knd == K_MULXOR)[:, None] | (knd == K_MULCOPY)[:, None],
                           _kara(cw, S), S)
            val = np.where((knd == K_ZERO)[:, None], 0, val)
            acc = (knd == K_XOR) | (knd == K_MULXOR)
            R[dst] = np.where(acc[:, None], R[dst] ^ val, val)
    return R[store]


def replay_lowered(trs, k, m, sym):
    steps, consts, ops, parity, regs = _lowered(trs, k, m)
    return replay(steps, consts, ops, np.arange(k), parity, regs, sym)


@pytest.mark.parametrize("k,m", REFERENCE_SHAPES + EXTRA_SHAPES)
def test_device_lowering_replays_to_matrix(trs, k, m):
    sym = _symbols(k * 7 + m, k, 6)
    assert (replay_lowered(trs, k, m, sym) == _matrix_parity(k, m, sym)).all()


def test_device_lowering_c3_shape(trs):
    # (1024,1024) against the matrix oracle on a few columns (the generator
    # build takes a few seconds in the C oracle).
    sym = _symbols(0xC3, 1024, 3)
    assert (replay_lowered(trs, 1024, 1024, sym) == _matrix_parity(1024, 1024, sym)).all()


# ---- additive-FFT erasure decoder ------------------------------------------

def _decode_program(trs, k, n, present):
    lib = trs._lib.lib
    flags = (ctypes.c_uint8 * n)(*[1 if p else 0 for p in present])
    counts = (ctypes.c_size_t * 6)()
    null = ctypes.POINTER(ctypes.c_uint32)()
    rc = lib.trs_fft16_decode_program(k, n, flags, null, 0, null, 0, null, 0, null, 0, null, 0,
                                      counts)
    assert rc == 0
    bufs = [np.zeros(max(1, int(c)), np.uint32) for c in counts[:5]]
    ptr = [b.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)) for b in bufs]
    rc = lib.trs_fft16_decode_program(k, n, flags, ptr[0], bufs[0].size, ptr[1], bufs[1].size,
                                      ptr[2], bufs[2].size, ptr[3], bufs[3].size, ptr[4],
                                      bufs[4].size, counts)
    assert rc == 0
    out = [b[:c].astype(np.int64) for b, c in zip(bufs, counts[:5])]
    return out[0].reshape(-1, 8), out[1], out[2], out[3], out[4], int(counts[5])


def _codeword(k, m, seed, cols):
    sym = _symbols(seed, k, cols)
    return np.concatenate([sym, _matrix_parity(k, m, sym)])


@pytest.mark.parametrize("k,m", [(2, 2), (4, 2), (3, 13), (7, 13), (10, 4), (10, 6), (17, 33),
                                 (32, 32), (100, 50), (200, 56), (300, 212), (1000, 24)])
def test_fft_decoder_replays_to_codeword(trs, k, m):
    """The decoder program rebuilds every erased symbol (data and parity) of
    oracle-encoded codewords, for random erasure patterns up to m shards."""
    n = k + m
    cw = _codeword(k, m, k * 31 + m, 5)
    rng = np.random.default_rng(k * 7 + m)
    for e in sorted({1, m, int(rng.integers(1, m + 1))}):
        erased = set(rng.choice(n, e, replace=False).tolist())
        present = [i not in erased for i in range(n)]
        steps, consts, ops, load, store, regs = _decode_program(trs, k, n, present)
        assert list(load) == [i for i in range(n) if present[i]][:k]
        assert list(store) == sorted(i for i in erased if i < k) + sorted(i for i in erased if i >= k)
        got = replay(steps, consts, ops, load, store, regs, cw[load])
        assert (got == cw[store]).all(), (k, m, sorted(erased))


def test_fft_decoder_c3_pattern(trs):
    # (1024,1024) with 1024 uniformly chosen erasures, a few symbol columns
    k = m = 1024
    cw = _codeword(k, m, 0xC3D, 2)
    rng = np.random.default_rng(0xC3)
    erased = set(rng.choice(k + m, 1024, replace=False).tolist())
    present = [i not in erased for i in range(k + m)]
    steps, consts, ops, load, store, regs = _decode_program(trs, k, k + m, present)
    assert regs == 2048
    # zero, scale by Lambda, 11 inverse levels, pre-scale, the formal derivative
    # as one step per popcount layer 0..10, post-scale, 11 forward levels,
    # scale by 1 / Lambda'
    assert len(steps) == 1 + 1 + 11 + 1 + 11 + 1 + 11 + 1
    got = replay(steps, consts, ops, load, store, regs, cw[load])
    assert (got == cw[store]).all()
