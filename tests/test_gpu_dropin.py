"""The drop-in host surface: every reference method of both codecs driven
from host buffers (numpy arrays, the reference's slices) through the C ABI's
*_host entry points, byte-compared with the oracle, plus the reference's
error order on host buffers (the same codes as the device calls), Clone and
the device-memory helpers for callers without a HIP runtime.

Reference methods (src/reedsolomon.rs, src/reedsolomon16.rs): encode (:271,
:412), encode_sep (:284, :422), encode_rows / encode_rows_into (:652, :667),
verify (:430, :453), reconstruct / reconstruct_data (:470, :508),
reconstruct_rows (:514, :546), prepare_decode + PreparedDecoder (:484, :518),
Clone (:725-739)."""
import ctypes

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _rows(k, length, seed):
    return O.fill(seed, k * length).reshape(k, length)


def _err(trs, fn, *a):
    with pytest.raises(trs.Error) as e:
        fn(*a)
    return e.value.name


# One shape per GF(2^8) encode route: generated bit-sliced FFT (7,13), the
# bit-sliced constant matrix (10,4), the staged FFT (48,16, up to 256 KiB) and
# the runtime-coefficient matrix kernels (20,10).
RS8_SHAPES = [(7, 13), (10, 4), (48, 16), (20, 10)]
LENGTHS = [1, 1000, 70_001, 300_000]


@pytest.fixture(autouse=True)
def small_chunks(trs):
    trs.set_host_chunk(64 << 10)  # several pipeline chunks per call
    yield
    trs.set_host_chunk(0)


@pytest.mark.parametrize("k,m", RS8_SHAPES)
def test_rs8_host_methods_match_oracle(trs, k, m):
    rs = trs.ReedSolomon(k, m)
    G = O.build_matrix(8, k, k + m)
    rng = np.random.default_rng(k * 100 + m)
    for length in LENGTHS:
        data = _rows(k, length, 31 + length)
        want = O.rs8_encode(G, k, data)
        full = np.concatenate([data, want])
        # encode over one shard list
        shards = [data[i].copy() for i in range(k)] + [np.zeros(length, np.uint8)
                                                      for _ in range(m)]
        rs.encode(shards)
        assert (np.stack(shards[k:]) == want).all(), (k, m, length)
        # encode_sep
        par = [np.full(length, 0xA5, np.uint8) for _ in range(m)]
        rs.encode_sep([data[i].copy() for i in range(k)], par)
        assert (np.stack(par) == want).all()
        # encode_rows / encode_rows_into over contiguous rows
        flat = np.ascontiguousarray(data).reshape(-1).copy()
        assert (rs.encode_rows(flat, length).reshape(m, length) == want).all()
        into = np.zeros(m * length, np.uint8)
        rs.encode_rows_into(flat, length, into)
        assert (into.reshape(m, length) == want).all()
        # verify: true, then single-bit flips at the start, middle and end of
        # a parity shard
        assert rs.verify([full[i].copy() for i in range(k + m)])
        for pos in (0, length // 2, length - 1):
            bad = [full[i].copy() for i in range(k + m)]
            bad[k + (pos % m)][pos] ^= 0x10
            assert not rs.verify(bad), (length, pos)
        # reconstruct / reconstruct_data over (buffer, present) pairs
        for data_only in (False, True):
            e = int(rng.integers(1, m + 1))
            erased = set(rng.choice(k + m, e, replace=False).tolist())
            rows = [full[i].copy() for i in range(k + m)]
            for i in erased:
                rows[i][:] = 0
            pairs = [(rows[i], i not in erased) for i in range(k + m)]
            (rs.reconstruct_data if data_only else rs.reconstruct)(pairs)
            for i in range(k + m):
                if data_only and i >= k and i in erased:
                    assert not rows[i].any()
                else:
                    assert (rows[i] == full[i]).all(), (length, sorted(erased), i, data_only)
        # Option<T> shards: None entries are allocated (host memory) and rebuilt
        opt = [full[i].copy() for i in range(k + m)]
        missing = sorted(rng.choice(k + m, m, replace=False).tolist())
        for i in missing:
            opt[i] = None
        rs.reconstruct(opt)
        assert all(isinstance(opt[i], np.ndarray) for i in missing)
        assert (np.stack(opt) == full).all()
        # reconstruct_rows
        erased = set(rng.choice(k + m, m, replace=False).tolist())
        rowbuf = full.copy()
        for i in erased:
            rowbuf[i] = 0
        flat = rowbuf.reshape(-1).copy()
        rs.reconstruct_rows(flat, length, [i not in erased for i in range(k + m)])
        assert (flat.reshape(k + m, length) == full).all()
        # prepared decoder: full and data-only, and a wrong pattern
        present = [i not in erased for i in range(k + m)]
        dec = rs.prepare_decode(present)
        for data_only in (False, True):
            rows = [full[i].copy() for i in range(k + m)]
            for i in erased:
                rows[i][:] = 0
            pairs = [(rows[i], present[i]) for i in range(k + m)]
            (dec.reconstruct_data if data_only else dec.reconstruct)(pairs)
            for i in range(k + m):
                if data_only and i >= k and i in erased:
                    assert not rows[i].any()
                else:
                    assert (rows[i] == full[i]).all()
        other = [not p for p in present]
        other[0], other[1] = True, True
        if sum(other) >= k and other != present:
            rows = [(full[i].copy(), other[i]) for i in range(k + m)]
            assert _err(trs, dec.reconstruct, rows) == "InvalidShardFlags"


@pytest.mark.parametrize("k,m,length", [(4, 2, 2), (4, 2, 4096), (17, 33, 70_002),
                                        (200, 100, 34), (1024, 1024, 64)])
def test_rs16_host_methods_match_oracle(trs, k, m, length):
    rs = trs.ReedSolomon16(k, m)
    G = O.build_matrix(16, k, k + m)
    rng = np.random.default_rng(k + m + length)
    data = _rows(k, length, 77 + k)
    want = O.rs16_symbolwise(G[k:], list(data), length)
    full = np.concatenate([data, want])
    shards = [data[i].copy() for i in range(k)] + [np.zeros(length, np.uint8) for _ in range(m)]
    rs.encode(shards)
    assert (np.stack(shards[k:]) == want).all()
    par = [np.zeros(length, np.uint8) for _ in range(m)]
    rs.encode_sep([data[i].copy() for i in range(k)], par)
    assert (np.stack(par) == want).all()
    assert rs.verify([full[i].copy() for i in range(k + m)])
    half = length // 2
    # flips in the low plane and the high plane of a symbol column
    for pos in (0, half - 1, half, length - 1):
        bad = [full[i].copy() for i in range(k + m)]
        bad[k + m - 1][pos] ^= 0x80
        assert not rs.verify(bad), pos
    for data_only in (False, True):
        e = min(m, int(rng.integers(1, m + 1)))
        erased = set(rng.choice(k + m, e, replace=False).tolist())
        rows = [full[i].copy() for i in range(k + m)]
        for i in erased:
            rows[i][:] = 0
        (rs.reconstruct_data if data_only else rs.reconstruct)(
            [(rows[i], i not in erased) for i in range(k + m)])
        for i in range(k + m):
            if data_only and i >= k and i in erased:
                assert not rows[i].any()
            else:
                assert (rows[i] == full[i]).all(), (sorted(erased)[:8], i, data_only)
    erased = set(rng.choice(k + m, m, replace=False).tolist())
    present = [i not in erased for i in range(k + m)]
    flat = full.copy()
    for i in erased:
        flat[i] = 0
    flat = flat.reshape(-1).copy()
    rs.reconstruct_rows(flat, length, present)
    assert (flat.reshape(k + m, length) == full).all()
    dec = rs.prepare_decode(present)
    rows = [full[i].copy() for i in range(k + m)]
    for i in erased:
        rows[i][:] = 0
    dec.reconstruct([(rows[i], present[i]) for i in range(k + m)])
    assert (np.stack(rows) == full).all()


def test_host_errors_follow_the_reference_order(trs):
    """Malformed host calls return the reference's errors, the same ones the
    device calls return (src/reedsolomon.rs:289-308, :701-722, :801-821,
    :520-528; src/reedsolomon16.rs:552-560, :758-772)."""
    import torch

    k, m, length = 4, 2, 64
    rs = trs.ReedSolomon(k, m)
    full = np.concatenate([_rows(k, length, 5), np.zeros((m, length), np.uint8)])
    host = [full[i].copy() for i in range(k + m)]
    dev = [torch.from_numpy(full[i].copy()).cuda() for i in range(k + m)]
    for shards in (host, dev):
        cast = (lambda a: a) if shards is host else (lambda a: torch.from_numpy(a).cuda())
        assert _err(trs, rs.encode, shards[:-1]) == "TooFewShards"
        assert _err(trs, rs.encode, shards + [shards[0]]) == "TooManyShards"
        assert _err(trs, rs.encode, [cast(np.zeros(0, np.uint8))] + shards[1:]) == "EmptyShard"
        ragged = shards[:-1] + [cast(np.zeros(length + 1, np.uint8))]
        assert _err(trs, rs.encode, ragged) == "IncorrectShardSize"
        assert _err(trs, rs.verify, ragged) == "IncorrectShardSize"
        assert _err(trs, rs.encode_sep, shards[:k - 1], shards[k:]) == "TooFewDataShards"
        assert _err(trs, rs.encode_sep, shards[:k], shards[k:] + [shards[0]]) == \
            "TooManyParityShards"
        few = [(s, i < k - 1) for i, s in enumerate(shards)]
        assert _err(trs, rs.reconstruct, few) == "TooFewShardsPresent"
        wrong = [(s, i != 0) for i, s in enumerate(shards)]
        wrong[0] = (cast(np.zeros(length - 1, np.uint8)), False)
        assert _err(trs, rs.reconstruct, wrong) == "IncorrectShardSize"
        rows = cast(np.zeros((k + m) * length, np.uint8))
        assert _err(trs, rs.reconstruct_rows, rows, 0, [True] * (k + m)) == "EmptyShard"
        assert _err(trs, rs.reconstruct_rows, rows, length, [True] * (k + m - 1)) == \
            "InvalidShardFlags"
        rs.reconstruct([(s, True) for s in shards])  # all present: a no-op
    rs16 = trs.ReedSolomon16(2, 2)
    odd = [np.zeros(5, np.uint8) for _ in range(4)]
    assert _err(trs, rs16.encode, odd) == "IncorrectShardSize"
    assert _err(trs, rs16.reconstruct_rows, np.zeros(20, np.uint8), 5, [True] * 4) == \
        "IncorrectShardSize"
    with pytest.raises(TypeError):  # host and device buffers do not mix in one call
        rs.encode(host[:k] + dev[k:])


def test_device_shard_in_a_host_call_is_rejected(trs):
    """A device pointer handed to a host entry point is an argument error."""
    import torch

    k, m, length = 4, 2, 256
    rs = trs.ReedSolomon(k, m)
    lib = trs._lib.lib
    host = [np.zeros(length, np.uint8) for _ in range(k + m)]
    d = torch.zeros(length, dtype=torch.uint8, device="cuda")
    ptrs = (ctypes.c_void_p * (k + m))(*([h.ctypes.data for h in host[:-1]] + [d.data_ptr()]))
    lens = (ctypes.c_size_t * (k + m))(*([length] * (k + m)))
    assert lib.trs_rs8_encode_host(rs._h, ptrs, lens, k + m) == -3


def test_clone_keeps_plans_and_shape(trs):
    import copy

    k, m, length = 10, 4, 4096
    rs = trs.ReedSolomon(k, m)
    G = O.build_matrix(8, k, k + m)
    data = _rows(k, length, 9)
    full = np.concatenate([data, O.rs8_encode(G, k, data)])
    present = [i not in (0, 5, 11) for i in range(k + m)]
    rs.prepare_decode(present)  # fills the plan cache
    c = copy.copy(rs)
    assert c == rs and c is not rs
    rows = [full[i].copy() for i in range(k + m)]
    for i in (0, 5, 11):
        rows[i][:] = 0
    c.reconstruct([(rows[i], present[i]) for i in range(k + m)])
    assert (np.stack(rows) == full).all()
    del rs  # the clone owns its own state
    assert c.verify([full[i].copy() for i in range(k + m)])
    c16 = trs.ReedSolomon16(4, 2).clone()
    assert c16.total_shard_count() == 6


def test_device_alloc_and_memcpy(trs):
    """trs_device_alloc / trs_memcpy / trs_device_free: a caller without a HIP
    runtime encodes through library-owned device buffers."""
    lib = trs._lib.lib
    k, m, length = 10, 4, 10_000
    rs = trs.ReedSolomon(k, m)
    G = O.build_matrix(8, k, k + m)
    data = _rows(k, length, 12)
    bufs = []
    for i in range(k + m):
        p = ctypes.c_void_p()
        assert lib.trs_device_alloc(length, ctypes.byref(p)) == 0
        bufs.append(p)
    try:
        for i in range(k):
            src = np.ascontiguousarray(data[i])
            assert lib.trs_memcpy(bufs[i], src.ctypes.data, length, None) == 0
        ptrs = (ctypes.c_void_p * (k + m))(*[b.value for b in bufs])
        lens = (ctypes.c_size_t * (k + m))(*([length] * (k + m)))
        assert lib.trs_rs8_encode(rs._h, ptrs, lens, k + m, None) == 0
        out = np.zeros((m, length), np.uint8)
        for j in range(m):
            assert lib.trs_memcpy(out[j].ctypes.data, bufs[k + j], length, None) == 0
        assert lib.trs_stream_synchronize(None) == 0
        assert (out == O.rs8_encode(G, k, data)).all()
    finally:
        for b in bufs:
            assert lib.trs_device_free(b) == 0
