"""Python view of the CPU oracle (oracle.c), for tests and the CPU baseline only.

TEST INFRASTRUCTURE.  Nothing in the shipped package imports this module; it is
the checker the GPU path is compared against.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use it.

Each wrapper names the reference function it restates.  How each one is pinned
(by something other than itself):

* GF(2^8) tables/mul/div/exp -- Backblaze LOG table and the known products of
  src/galois.rs:119-151 (golden fixture), field axioms, and an independent
  carry-less polynomial multiply in tests/test_oracle.py.
* gfni_matrix -- identity qword 0x0102040810204080 (src/gf/gfni.rs:85-86) and
  the affine model == mul for all 65,536 pairs (src/gf/gfni.rs:99-107).
* GF(2^16) mul/inv/const_matrix -- TOWER_L minimality and const_matrix(1),
  const_matrix(0x100) (src/galois16.rs:143-196), inverse of every element,
  associativity/distributivity samples, and an independent polynomial-basis
  multiply modulo y^2+y+32.
* invert/vandermonde/build_matrix -- V*V^-1 == I (src/matrix.rs:205-225),
  singular rejection (:229-236), systematic top block == I, MDS (every k-row
  subset invertible) on small shapes, and an independent numpy restatement.
* rs8 encode/reconstruct -- the generator pins above plus round trips (erase ->
  reconstruct == original) across the reference test patterns.
* rs16 apply -- RS16(3,2) frozen wire vector (src/reedsolomon16.rs:1087-1101)
  and the symbol-wise tower oracle (src/reedsolomon16.rs:919-936).
"""
from __future__ import annotations

import ctypes
import os
from typing import Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_u8p = ctypes.POINTER(ctypes.c_uint8)
_u16p = ctypes.POINTER(ctypes.c_uint16)


def build() -> None:
    """Compile liboracle.so with the Makefile beside this file."""
    import subprocess

    subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_gf8_mul.restype = ctypes.c_uint8
        L.orc_gf8_mul.argtypes = [ctypes.c_uint8, ctypes.c_uint8]
        L.orc_gf8_div.restype = ctypes.c_uint8
        L.orc_gf8_div.argtypes = [ctypes.c_uint8, ctypes.c_uint8]
        L.orc_gf8_exp.restype = ctypes.c_uint8
        L.orc_gf8_exp.argtypes = [ctypes.c_uint8, ctypes.c_size_t]
        L.orc_gfni_matrix.restype = ctypes.c_uint64
        L.orc_gfni_matrix.argtypes = [ctypes.c_uint8]
        L.orc_gf16_mul.restype = ctypes.c_uint16
        L.orc_gf16_mul.argtypes = [ctypes.c_uint16, ctypes.c_uint16]
        L.orc_gf16_mul_fast.restype = ctypes.c_uint16
        L.orc_gf16_mul_fast.argtypes = [ctypes.c_uint16, ctypes.c_uint16]
        L.orc_gf16_inv.restype = ctypes.c_uint16
        L.orc_gf16_inv.argtypes = [ctypes.c_uint16]
        L.orc_gf16_const_matrix.argtypes = [ctypes.c_uint16, _u8p]
        L.orc_gf8_tables.argtypes = [_u8p, _u8p]
        L.orc_vandermonde.argtypes = [ctypes.c_int, _u16p, ctypes.c_size_t, ctypes.c_size_t]
        L.orc_multiply.argtypes = [ctypes.c_int, _u16p, _u16p, _u16p,
                                   ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t]
        L.orc_invert.restype = ctypes.c_int
        L.orc_invert.argtypes = [ctypes.c_int, _u16p, _u16p, ctypes.c_size_t]
        L.orc_build_matrix.restype = ctypes.c_int
        L.orc_build_matrix.argtypes = [ctypes.c_int, ctypes.c_size_t, ctypes.c_size_t, _u16p]
        L.orc_mul_slice.argtypes = [_u8p, _u8p, ctypes.c_size_t, ctypes.c_uint8, ctypes.c_int]
        pp = ctypes.POINTER(_u8p)
        L.orc_rs8_apply.argtypes = [_u8p, ctypes.c_size_t, ctypes.c_size_t, pp, pp, ctypes.c_size_t]
        L.orc_rs8_reconstruct.restype = ctypes.c_int
        L.orc_rs8_reconstruct.argtypes = [_u8p, ctypes.c_size_t, ctypes.c_size_t, pp, _u8p,
                                          ctypes.c_size_t, ctypes.c_int]
        for name in ("orc_rs16_apply", "orc_rs16_symbolwise"):
            getattr(L, name).argtypes = [_u16p, ctypes.c_size_t, ctypes.c_size_t, pp, pp,
                                         ctypes.c_size_t]
        L.orc_rs16_apply_strips.argtypes = [_u16p, ctypes.c_size_t, ctypes.c_size_t, pp, pp,
                                            ctypes.c_size_t, ctypes.c_size_t]
        L.orc_rs16_reconstruct.restype = ctypes.c_int
        L.orc_rs16_reconstruct.argtypes = [_u16p, ctypes.c_size_t, ctypes.c_size_t, pp, _u8p,
                                           ctypes.c_size_t, ctypes.c_int, ctypes.c_size_t]
        L.orc_fill.argtypes = [ctypes.c_uint64, _u8p, ctypes.c_size_t]
        L.orc_simd_available.restype = ctypes.c_int
        L.orc_simd_apply.argtypes = [_u8p, ctypes.c_size_t, ctypes.c_size_t, pp, pp,
                                     ctypes.c_size_t]
        _i32p = ctypes.POINTER(ctypes.c_int32)
        L.orc_simd_fft_encode.argtypes = [_i32p, ctypes.c_size_t, ctypes.c_size_t,
                                          ctypes.c_size_t, ctypes.c_size_t, _i32p, pp, pp,
                                          ctypes.c_size_t]
        L.orc_simd_fft_generated.restype = ctypes.c_int
        L.orc_simd_fft_generated.argtypes = [ctypes.c_size_t, ctypes.c_size_t, pp, pp,
                                             ctypes.c_size_t]
        L.orc_simd_rs8_reconstruct.restype = ctypes.c_int
        L.orc_simd_rs8_reconstruct.argtypes = [_u8p, ctypes.c_size_t, ctypes.c_size_t, pp, _u8p,
                                               ctypes.c_size_t, ctypes.c_int]
        _lib = L
    return _lib


def _p8(a: np.ndarray):
    assert a.dtype == np.uint8 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_u8p)


def _p16(a: np.ndarray):
    assert a.dtype == np.uint16 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_u16p)


def _ptr_array(rows: Sequence[np.ndarray]):
    arr = (_u8p * max(1, len(rows)))()
    for i, r in enumerate(rows):
        assert r.dtype == np.uint8 and r.flags["C_CONTIGUOUS"]
        arr[i] = r.ctypes.data_as(_u8p)
    return arr


# -- GF(2^8): src/galois.rs ------------------------------------------------

def gf8_tables():
    """LOG (256) and doubled EXP (510) tables, src/galois.rs:9-35."""
    log = np.zeros(256, np.uint8)
    exp = np.zeros(510, np.uint8)
    lib().orc_gf8_tables(_p8(log), _p8(exp))
    return log, exp


def gf8_mul(a: int, b: int) -> int:
    return lib().orc_gf8_mul(a, b)


def gf8_div(a: int, b: int) -> int:
    return lib().orc_gf8_div(a, b)


def gf8_exp(a: int, n: int) -> int:
    return lib().orc_gf8_exp(a, n)


def gf8_mul_table() -> np.ndarray:
    log, exp = gf8_tables()
    a = np.arange(256)
    t = exp[(log[a][:, None].astype(np.int64) + log[a][None, :])]
    t[0, :] = 0
    t[:, 0] = 0
    return t.astype(np.uint8)


def gfni_matrix(c: int) -> int:
    """src/gf/gfni.rs:47-64."""
    return lib().orc_gfni_matrix(c)


# -- GF((2^8)^2): src/galois16.rs ------------------------------------------

def gf16_mul(a: int, b: int) -> int:
    return lib().orc_gf16_mul(a, b)


def gf16_mul_fast(a: int, b: int) -> int:
    return lib().orc_gf16_mul_fast(a, b)


def gf16_inv(a: int) -> int:
    return lib().orc_gf16_inv(a)


def gf16_const_matrix(c: int):
    out = np.zeros(4, np.uint8)
    lib().orc_gf16_const_matrix(c, _p8(out))
    return [int(x) for x in out]


# -- matrices: src/matrix.rs -------------------------------------------------

def _dt(field: int):
    return np.uint8 if field == 8 else np.uint16


def vandermonde(field: int, rows: int, cols: int) -> np.ndarray:
    out = np.zeros((rows, cols), np.uint16)
    lib().orc_vandermonde(field, _p16(out), rows, cols)
    return out.astype(_dt(field))


def multiply(field: int, a: np.ndarray, b: np.ndarray) -> np.ndarray:
    a16 = np.ascontiguousarray(a, dtype=np.uint16)
    b16 = np.ascontiguousarray(b, dtype=np.uint16)
    out = np.zeros((a.shape[0], b.shape[1]), np.uint16)
    lib().orc_multiply(field, _p16(a16), _p16(b16), _p16(out), a.shape[0], a.shape[1], b.shape[1])
    return out.astype(_dt(field))


def invert(field: int, m: np.ndarray):
    """Returns (status, inverse); status 14 = SingularMatrix."""
    n = m.shape[0]
    m16 = np.ascontiguousarray(m, dtype=np.uint16)
    out = np.zeros((n, n), np.uint16)
    rc = lib().orc_invert(field, _p16(m16), _p16(out), n)
    return rc, out.astype(_dt(field))


def build_matrix(field: int, k: int, n: int) -> np.ndarray:
    """Systematic generator V(n,k)*inv(V_top), src/reedsolomon.rs:154-159."""
    out = np.zeros((n, k), np.uint16)
    rc = lib().orc_build_matrix(field, k, n, _p16(out))
    assert rc == 0
    return out.astype(_dt(field))


# -- codec operations ----------------------------------------------------------

def mul_slice(out: np.ndarray, inp: np.ndarray, c: int, xor: bool) -> None:
    """src/gf/scalar.rs:7-23."""
    assert out.shape == inp.shape
    lib().orc_mul_slice(_p8(out), _p8(inp), out.size, c, int(xor))


def rs8_apply(rows: np.ndarray, inputs: Sequence[np.ndarray], length: int) -> np.ndarray:
    """RowTables::apply, src/gf/tables.rs:162-219."""
    rows = np.ascontiguousarray(rows, dtype=np.uint8)
    nout, k = rows.shape
    out = np.zeros((nout, length), np.uint8)
    outs = [out[j] for j in range(nout)]
    lib().orc_rs8_apply(_p8(rows), nout, k, _ptr_array(list(inputs)), _ptr_array(outs), length)
    return out


def rs8_encode(G: np.ndarray, k: int, data: np.ndarray) -> np.ndarray:
    """Parity rows of encode (src/reedsolomon.rs:271-351); data is (k, L)."""
    data = np.ascontiguousarray(data, dtype=np.uint8)
    return rs8_apply(G[k:], [data[i] for i in range(k)], data.shape[1])


def rs8_reconstruct(G: np.ndarray, k: int, m: int, shards: np.ndarray,
                    present: Sequence[bool], data_only: bool = False) -> int:
    """In-place reconstruct of an (n, L) array, src/reedsolomon.rs:550-891."""
    G = np.ascontiguousarray(G, dtype=np.uint8)
    pres = np.array([1 if p else 0 for p in present], np.uint8)
    rows = [shards[i] for i in range(k + m)]
    return lib().orc_rs8_reconstruct(_p8(G), k, m, _ptr_array(rows), _p8(pres),
                                     shards.shape[1], int(data_only))


def rs16_apply(rows: np.ndarray, inputs: Sequence[np.ndarray], length: int,
               strip_bytes: int = 0) -> np.ndarray:
    """RowTables16::apply (src/reedsolomon16.rs:143-168), optionally strip-mined
    the way encode_into does it (src/reedsolomon16.rs:364-385)."""
    rows = np.ascontiguousarray(rows, dtype=np.uint16)
    nout, k = rows.shape
    out = np.zeros((nout, length), np.uint8)
    outs = [out[j] for j in range(nout)]
    lib().orc_rs16_apply_strips(_p16(rows), nout, k, _ptr_array(list(inputs)),
                                _ptr_array(outs), length, strip_bytes)
    return out


def rs16_symbolwise(rows: np.ndarray, inputs: Sequence[np.ndarray], length: int) -> np.ndarray:
    """Symbol-wise tower parity, src/reedsolomon16.rs:919-936."""
    rows = np.ascontiguousarray(rows, dtype=np.uint16)
    nout, k = rows.shape
    out = np.zeros((nout, length), np.uint8)
    outs = [out[j] for j in range(nout)]
    lib().orc_rs16_symbolwise(_p16(rows), nout, k, _ptr_array(list(inputs)), _ptr_array(outs),
                              length)
    return out


def rs16_reconstruct(G: np.ndarray, k: int, m: int, shards: np.ndarray,
                     present: Sequence[bool], data_only: bool = False,
                     strip_bytes: int = 0) -> int:
    G = np.ascontiguousarray(G, dtype=np.uint16)
    pres = np.array([1 if p else 0 for p in present], np.uint8)
    rows = [shards[i] for i in range(k + m)]
    return lib().orc_rs16_reconstruct(_p16(G), k, m, _ptr_array(rows), _p8(pres),
                                      shards.shape[1], int(data_only), strip_bytes)


def fill(seed: int, length: int) -> np.ndarray:
    """Counter-based splitmix64 byte stream (the seeded-input generator)."""
    out = np.zeros(length, np.uint8)
    lib().orc_fill(seed & 0xFFFFFFFFFFFFFFFF, _p8(out), length)
    return out


# -- x86 GFNI / AVX-512 restatement (simd_ref.c): the CPU baseline -----------
# Pinned byte for byte against the scalar functions above (tests/test_simd_ref.py).

_FFT_KINDS = {"xo": 0, "cp": 1, "mx": 2, "mc": 3, "zero": 4}


def simd_available() -> bool:
    """True when this CPU has GFNI, AVX-512F/BW/VL (the simd_* functions need them)."""
    return bool(lib().orc_simd_available())


def simd_rs8_apply(rows: np.ndarray, inputs: Sequence[np.ndarray], length: int) -> np.ndarray:
    """RowTables::apply on the fused GFNI kernel (src/gf/x86_fused.rs:13-130)."""
    rows = np.ascontiguousarray(rows, dtype=np.uint8)
    nout, k = rows.shape
    out = np.zeros((nout, length), np.uint8)
    lib().orc_simd_apply(_p8(rows), nout, k, _ptr_array(list(inputs)),
                         _ptr_array([out[j] for j in range(nout)]), length)
    return out


def simd_fft_program(ops, parity_regs, register_count: int):
    """Packs an LCH op list ((kind, dst, src[, c]) tuples, fftgen's format) for
    simd_fft_encode."""
    flat = []
    for op in ops:
        kind = _FFT_KINDS[op[0]]
        flat += [kind, op[1], op[2] if len(op) > 2 else 0, op[3] if len(op) > 3 else 0]
    return (np.array(flat, np.int32), np.array(list(parity_regs), np.int32), int(register_count))


def simd_fft_encode(prog, k: int, data: np.ndarray) -> np.ndarray:
    """Parity of `data` (k, L) through a packed LCH program (src/gf/fft_x86.rs:22-165)."""
    ops, preg, nreg = prog
    data = np.ascontiguousarray(data, dtype=np.uint8)
    m, length = len(preg), data.shape[1]
    out = np.zeros((m, length), np.uint8)
    i32 = ctypes.POINTER(ctypes.c_int32)
    lib().orc_simd_fft_encode(ops.ctypes.data_as(i32), len(ops) // 4, nreg, k, m,
                              preg.ctypes.data_as(i32), _ptr_array([data[i] for i in range(k)]),
                              _ptr_array([out[j] for j in range(m)]), length)
    return out


def simd_rs8_reconstruct(G: np.ndarray, k: int, m: int, shards: np.ndarray,
                         present: Sequence[bool], data_only: bool = False) -> int:
    """In-place reconstruct on the fused GFNI kernel (src/reedsolomon.rs:829-891)."""
    G = np.ascontiguousarray(G, dtype=np.uint8)
    pres = np.array([1 if p else 0 for p in present], np.uint8)
    rows = [shards[i] for i in range(k + m)]
    return lib().orc_simd_rs8_reconstruct(_p8(G), k, m, _ptr_array(rows), _p8(pres),
                                          shards.shape[1], int(data_only))


def simd_fft_generated(k: int, m: int, data: np.ndarray, out: np.ndarray = None):
    """Parity through the straight-line encoder of a generated shape
    (simd_fft_gen.c) into `out` (allocated when None); None when (k, m) has no
    generated program."""
    data = np.ascontiguousarray(data, dtype=np.uint8)
    length = data.shape[1]
    if out is None:
        out = np.zeros((m, length), np.uint8)
    rc = lib().orc_simd_fft_generated(k, m, _ptr_array([data[i] for i in range(k)]),
                                      _ptr_array([out[j] for j in range(m)]), length)
    return None if rc else out


# Prepared calls for timing: pointer arrays are built once, so a timed call is
# one C call (building ctypes arrays costs microseconds per row).

def simd_encode_call(G: np.ndarray, k: int, m: int, data: np.ndarray, parity: np.ndarray):
    """A zero-argument callable writing the parity of `data` into `parity`:
    the straight-line program for a generated shape (the reference's
    fft-generated route), the fused GFNI apply of G's parity rows otherwise
    (src/reedsolomon.rs:204-234 routes the same way)."""
    length = data.shape[1]
    ins = _ptr_array([data[i] for i in range(k)])
    outs = _ptr_array([parity[j] for j in range(m)])
    L = lib()
    # the callables hold the arrays their pointer tables point into
    if L.orc_simd_fft_generated(k, m, None, None, 0) == 0:  # the shape has a program
        return lambda keep=(data, parity): L.orc_simd_fft_generated(k, m, ins, outs, length)
    rows = np.ascontiguousarray(G[k:], dtype=np.uint8)
    rp = _p8(rows)
    return lambda keep=(rows, data, parity): L.orc_simd_apply(rp, m, k, ins, outs, length)


def rs8_decode_plan(G: np.ndarray, k: int, m: int, present: Sequence[bool]):
    """DecodePlan (src/reedsolomon.rs:537-613): the first k present shards,
    rows rebuilding each missing data shard from them, and the missing parity
    indices (rebuilt from the data with G's rows)."""
    valid = [i for i in range(k + m) if present[i]][:k]
    sub = np.array([G[i] for i in valid], np.uint16)
    rc, inv = invert(8, sub)
    assert rc == 0
    mdata = [i for i in range(k) if not present[i]]
    mpar = [i for i in range(k, k + m) if not present[i]]
    drows = np.array([inv[i] for i in mdata], np.uint8).reshape(len(mdata), k)
    return valid, mdata, drows, mpar


def simd_reconstruct_calls(G: np.ndarray, k: int, m: int, shards: np.ndarray, plan):
    """Callables for the two fused applies of reconstruct_with_plan
    (src/reedsolomon.rs:829-891): missing data from the plan's survivors, then
    missing parity from the data."""
    valid, mdata, drows, mpar = plan
    length = shards.shape[1]
    L = lib()
    calls = []
    if mdata:
        src = _ptr_array([shards[i] for i in valid])
        dst = _ptr_array([shards[i] for i in mdata])
        dp = _p8(drows)
        calls.append(lambda keep=(drows, shards): L.orc_simd_apply(dp, len(mdata), k, src, dst,
                                                                  length))
    if mpar:
        prows = np.ascontiguousarray(G[mpar], dtype=np.uint8)
        src2 = _ptr_array([shards[i] for i in range(k)])
        dst2 = _ptr_array([shards[i] for i in mpar])
        pp2 = _p8(prows)
        calls.append(lambda keep=(prows, shards): L.orc_simd_apply(pp2, len(mpar), k, src2, dst2,
                                                                  length))
    return calls
