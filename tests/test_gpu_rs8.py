"""GPU parity for ReedSolomon (GF(2^8)) through the C ABI, against the oracle.

Mirrors the reference's codec tests (src/reedsolomon.rs:917-1207) with the
CPU oracle as the expected side, plus the error order of the boundary and the
full-size (10,4) x 256 x 64 KiB workload (round trip + oracle parity)."""
import numpy as np
import pytest

import oracle as O
from util import deterministic_data, rows_to_dev, seeded, to_dev, to_host

pytestmark = pytest.mark.gpu

LENGTHS = [1, 15, 16, 17, 31, 32, 63, 100, 1024, 10000, 65536]


def encoded(trs, k, m, length, seed=0x1234_5678):
    rs = trs.ReedSolomon(k, m)
    data = deterministic_data(k * length, seed).reshape(k, length) if length * k <= 200_000 \
        else seeded(k, length, seed)
    shards = rows_to_dev(np.concatenate([data, np.zeros((m, length), np.uint8)]))
    rs.encode(shards)
    return rs, shards, data


@pytest.mark.parametrize("k,m", [(7, 13), (10, 10), (20, 10), (4, 2), (6, 4), (17, 17), (10, 4)])
def test_encode_matches_oracle(trs, k, m):
    # src/reedsolomon.rs:944-970 (routes agree with encode_scalar across lengths)
    G = O.build_matrix(8, k, k + m)
    assert (trs.ReedSolomon(k, m).generator() == G).all()
    for length in LENGTHS:
        rs, shards, data = encoded(trs, k, m, length, 0x9E37_79B9)
        want = O.rs8_encode(G, k, data)
        assert (to_host(shards[k:]) == want).all(), (k, m, length)
        assert (to_host(shards[:k]) == data).all()


def test_wide_shapes_and_output_tiles(trs):
    # k + m up to 256; more than 16 outputs exercises the output-tile grid
    for k, m in [(1, 1), (3, 40), (128, 128), (255, 1), (1, 255)]:
        G = O.build_matrix(8, k, k + m)
        for length in (1, 33, 4096 + 5):
            data = seeded(k, length, 77 + k)
            rs = trs.ReedSolomon(k, m)
            shards = rows_to_dev(np.concatenate([data, np.zeros((m, length), np.uint8)]))
            rs.encode(shards)
            assert (to_host(shards[k:]) == O.rs8_encode(G, k, data)).all(), (k, m, length)


def test_encode_rows_matches_per_shard_encode(trs):
    # src/reedsolomon.rs:917-942
    k, m, row_len = 6, 4, 293
    rs, shards, data = encoded(trs, k, m, row_len)
    expected = np.concatenate(to_host(shards[k:]))
    flat = to_dev(data.reshape(-1))
    assert (rs.encode_rows(flat, row_len).cpu().numpy() == expected).all()
    buf = to_dev(np.zeros(m * row_len, np.uint8))
    rs.encode_rows_into(flat, row_len, buf)
    assert (buf.cpu().numpy() == expected).all()
    with pytest.raises(trs.Error) as e:
        rs.encode_rows_into(flat, row_len + 1, buf)
    assert e.value.name == "IncorrectShardSize"


def test_encode_sep_matches_encode(trs):
    # src/reedsolomon.rs:974-993
    for k, m in [(7, 13), (10, 10), (32, 32), (4, 2), (18, 6)]:
        for length in (16, 100, 987, 1024):
            rs, shards, data = encoded(trs, k, m, length)
            parity = rows_to_dev(np.zeros((m, length), np.uint8))
            rs.encode_sep(rows_to_dev(data), parity)
            assert (to_host(parity) == to_host(shards[k:])).all()


def test_encode_fused_alias(trs):
    rs, shards, data = encoded(trs, 10, 10, 1000)
    again = rows_to_dev(np.concatenate([data, np.zeros((10, 1000), np.uint8)]))
    rs.encode_fused(again)
    assert (to_host(again) == to_host(shards)).all()


def _erase_and_view(shards, erased):
    view = []
    for i, s in enumerate(shards):
        if i in erased:
            s.zero_()
        view.append((s, i not in erased))
    return view


@pytest.mark.parametrize("k,m", [(4, 2), (7, 13), (10, 10), (6, 4), (10, 4), (32, 32)])
@pytest.mark.parametrize("length", [15, 100, 1024])
def test_reconstruct_patterns(trs, k, m, length):
    # src/reedsolomon.rs:1048-1084, each pattern twice (second hits the plan cache)
    rs, shards, data = encoded(trs, k, m, length)
    original = to_host(shards)
    patterns = [[0], [k - 1], [k], list(range(min(m, k))), list(range(k, k + m)),
                [(e * 2) % (k + m) for e in range(m)]]
    for erased in patterns:
        for _ in range(2):
            cur = rows_to_dev(original)
            rs.reconstruct(_erase_and_view(cur, set(erased)))
            assert (to_host(cur) == original).all(), (k, m, length, erased)


def test_round_trip_reconstruct(trs):
    # src/reedsolomon.rs:1023-1045
    rs = trs.ReedSolomon(4, 4)
    data = np.stack([np.full(100, j + 1, np.uint8) for j in range(4)])
    shards = rows_to_dev(np.concatenate([data, np.zeros((4, 100), np.uint8)]))
    rs.encode(shards)
    original = to_host(shards)
    rs.reconstruct(_erase_and_view(shards, {0, 2, 5}))
    assert (to_host(shards) == original).all()


def test_option_shards_are_allocated(trs):
    # Option<T> containers: None is allocated zeroed and filled (src/reedsolomon.rs:79-102)
    k, m, length = 5, 3, 300
    rs, shards, data = encoded(trs, k, m, length)
    original = to_host(shards)
    opt = list(shards)
    opt[1] = None
    opt[6] = None
    rs.reconstruct(opt)
    assert all(t is not None for t in opt)
    assert (to_host(opt) == original).all()
    opt = list(rows_to_dev(original))
    opt[0] = None
    opt[7] = None
    rs.reconstruct_data(opt)
    assert (opt[0].cpu().numpy() == original[0]).all() and opt[7] is None


def test_prepared_decoder(trs):
    # src/reedsolomon.rs:1087-1129
    k, m, length = 7, 13, 200
    rs, shards, data = encoded(trs, k, m, length)
    original = to_host(shards)
    erased = {1, 3, 8, 15}
    dec = rs.prepare_decode([i not in erased for i in range(k + m)])
    for _ in range(3):
        cur = rows_to_dev(original)
        dec.reconstruct(_erase_and_view(cur, erased))
        assert (to_host(cur) == original).all()
    cur = rows_to_dev(original)
    with pytest.raises(trs.Error) as e:
        dec.reconstruct(_erase_and_view(cur, {0}))
    assert e.value.name == "InvalidShardFlags"


def test_prepared_data_only(trs):
    # src/reedsolomon.rs:1132-1160
    k, m, length = 6, 4, 128
    rs, shards, data = encoded(trs, k, m, length)
    original = to_host(shards)
    erased = {2, k + 1}
    dec = rs.prepare_decode([i not in erased for i in range(k + m)])
    cur = rows_to_dev(original)
    dec.reconstruct_data(_erase_and_view(cur, erased))
    got = to_host(cur)
    assert (got[2] == original[2]).all()
    assert (got[k + 1] == 0).all(), "missing parity must stay untouched"


def test_reconstruct_rows_roundtrip(trs):
    # src/reedsolomon.rs:1163-1185 (row_len 257: unaligned rows take the byte path)
    k, m, row_len = 10, 10, 257
    rs, shards, data = encoded(trs, k, m, row_len)
    pristine = to_host(shards).reshape(-1)
    rows = pristine.copy()
    erased = [0, 4, 12, 19]
    present = [True] * (k + m)
    for e in erased:
        present[e] = False
        rows[e * row_len:(e + 1) * row_len] = 0
    drows = to_dev(rows)
    rs.reconstruct_rows(drows, row_len, present)
    assert (drows.cpu().numpy() == pristine).all()
    with pytest.raises(trs.Error) as e:
        rs.reconstruct_rows(drows, row_len, present[:-1])
    assert e.value.name == "InvalidShardFlags"
    with pytest.raises(trs.Error) as e:
        rs.reconstruct_rows(drows, 0, present)
    assert e.value.name == "EmptyShard"
    with pytest.raises(trs.Error) as e:
        rs.reconstruct_rows(drows, row_len - 1, present)
    assert e.value.name == "IncorrectShardSize"


def test_verify_blocks(trs):
    # src/reedsolomon.rs:1188-1207
    k, m = 4, 3
    # the last two lengths take the bit-sliced planes kernel (from 128 KiB)
    for length in (100, 32 * 1024, 32 * 1024 + 17, 256 * 1024, 256 * 1024 + 17):
        rs, shards, data = encoded(trs, k, m, length)
        assert rs.verify(shards)
        host = to_host(shards)
        for pos in (0, length // 2, length - 1):
            bad = host.copy()
            bad[k, pos] ^= 1
            assert not rs.verify(rows_to_dev(bad)), (length, pos)


def test_error_order(trs):
    # src/reedsolomon.rs:701-722 (encode) and :615-641, :801-821 (reconstruct)
    k, m, length = 4, 2, 64
    rs, shards, data = encoded(trs, k, m, length)

    def err(fn, *a):
        with pytest.raises(trs.Error) as e:
            fn(*a)
        return e.value.name

    assert err(rs.encode, shards[:-1]) == "TooFewShards"
    assert err(rs.encode, shards + [shards[0]]) == "TooManyShards"
    empty = [to_dev(np.zeros(0, np.uint8))] + shards[1:]
    assert err(rs.encode, empty) == "EmptyShard"
    ragged = shards[:-1] + [to_dev(np.zeros(length + 1, np.uint8))]
    assert err(rs.encode, ragged) == "IncorrectShardSize"
    assert err(rs.verify, ragged) == "IncorrectShardSize"
    assert err(rs.encode_sep, shards[:k - 1], shards[k:]) == "TooFewDataShards"
    assert err(rs.encode_sep, shards[:k] + [shards[0]], shards[k:]) == "TooManyDataShards"
    assert err(rs.encode_sep, shards[:k], shards[k:k + 1]) == "TooFewParityShards"
    assert err(rs.encode_sep, shards[:k], shards[k:] + [shards[0]]) == "TooManyParityShards"
    assert err(rs.reconstruct, [(s, True) for s in shards[:-1]]) == "TooFewShards"
    few = [(s, i < k - 1) for i, s in enumerate(shards)]
    assert err(rs.reconstruct, few) == "TooFewShardsPresent"
    wrong_missing = [(s, i != 0) for i, s in enumerate(shards)]
    wrong_missing[0] = (to_dev(np.zeros(length - 1, np.uint8)), False)
    assert err(rs.reconstruct, wrong_missing) == "IncorrectShardSize"
    assert err(rs.prepare_decode, [True] * (k + m - 1)) == "InvalidShardFlags"
    assert err(rs.prepare_decode, [True] * (k - 1) + [False] * (m + 1)) == "TooFewShardsPresent"
    # all present: a no-op
    rs.reconstruct([(s, True) for s in shards])


def test_mul_slice_matches_scalar(trs):
    # src/gf/scalar.rs:7-23 via src/gf/mod.rs:69-85
    for length in (1, 7, 16, 4097, 100_003):
        inp = O.fill(length, length)
        base = O.fill(length + 1, length)
        for c in (0, 1, 2, 0x1d, 0x8e, 255):
            out = to_dev(base)
            trs.mul_slice(out, to_dev(inp), c)
            want = base.copy()
            O.mul_slice(want, inp, c, False)
            assert (out.cpu().numpy() == want).all()
            out = to_dev(base)
            trs.mul_slice_xor(out, to_dev(inp), c)
            want = base.copy()
            O.mul_slice(want, inp, c, True)
            assert (out.cpu().numpy() == want).all()
    with pytest.raises(AssertionError):
        trs.mul_slice(to_dev(np.zeros(3, np.uint8)), to_dev(np.zeros(4, np.uint8)), 3)


def test_fill_random_matches_oracle_stream(trs):
    import torch

    for length in (1, 8, 13, 4096, 100_001):
        t = torch.empty(length, dtype=torch.uint8, device="cuda")
        trs.fill_random(t, 0x7A9E5EED00000001 + length)
        assert (t.cpu().numpy() == O.fill(0x7A9E5EED00000001 + length, length)).all()


def test_full_size_10_4_workload(trs):
    """256 stripes x (10,4) x 64 KiB, seeded: parity == oracle for every stripe,
    then 4 uniformly chosen erasures per stripe reconstruct to the original."""
    import torch

    k, m, L, S = 10, 4, 65536, 256
    rs = trs.ReedSolomon(k, m)
    G = O.build_matrix(8, k, k + m)
    buf = torch.empty(S * (k + m) * L, dtype=torch.uint8, device="cuda")
    rng = np.random.default_rng(0x7A9E5EED00000000)
    for s in range(S):
        stripe = buf[s * (k + m) * L:(s + 1) * (k + m) * L]
        trs.fill_random(stripe[:k * L], 0x7A9E5EED00000000 + s)
        rs.encode([stripe[i * L:(i + 1) * L] for i in range(k + m)])
    host = buf.cpu().numpy().reshape(S, k + m, L)
    for s in range(S):
        data = O.fill(0x7A9E5EED00000000 + s, k * L).reshape(k, L)
        assert (host[s, :k] == data).all()
        assert (host[s, k:] == O.rs8_encode(G, k, data)).all(), s
    original = buf.clone()
    for s in range(S):
        erased = set(rng.choice(k + m, 4, replace=False).tolist())
        stripe = buf[s * (k + m) * L:(s + 1) * (k + m) * L]
        shards = [stripe[i * L:(i + 1) * L] for i in range(k + m)]
        rs.reconstruct(_erase_and_view(shards, erased))
    torch.cuda.synchronize()
    assert torch.equal(buf, original)


@pytest.mark.parametrize("k,m", [(7, 13), (10, 10), (14, 14), (16, 16), (18, 6), (32, 32),
                                 (10, 4)])
def test_every_encode_route_matches_oracle(trs, k, m):
    """Table kernel, bit-sliced constant matrix and bit-sliced LCH FFT are
    byte-identical to the oracle across tails and odd lengths."""
    G = O.build_matrix(8, k, k + m)
    rs = trs.ReedSolomon(k, m)
    for length in (1, 31, 32, 33, 1000, 8192, 8192 + 17, 65536):
        data = seeded(k, length, 4242 + length)
        want = O.rs8_encode(G, k, data)
        for route in ("table", "bitsliced-matrix", "bitsliced-fft"):
            parity = rows_to_dev(np.zeros((m, length), np.uint8))
            try:
                rs.encode_sep_route(route, rows_to_dev(data), parity)
            except trs.Error as e:
                assert e.code == -4 and route != "table"
                continue
            assert (to_host(parity) == want).all(), (k, m, length, route)
    assert rs.encode_route(65536).startswith("gpu-")


def test_pair_program_per_call_unaligned(trs):
    """The per-call (32,32) encode and verify kernels (one wave pair per
    2 KiB of shard) on shards that start one byte past an allocation (byte
    accesses), with lengths that end inside different 2 KiB windows of a
    tile; verify catches a flipped last byte."""
    import torch

    k, m = 32, 32
    G = O.build_matrix(8, k, k + m)
    rs = trs.ReedSolomon(k, m)
    for length in (2048 + 16, 4096 + 5, 6144 + 1, 8192 + 2048 + 3):
        data = seeded(k, length, 77 + length)
        want = O.rs8_encode(G, k, data)
        bufs = [torch.zeros(length + 1, dtype=torch.uint8, device="cuda") for _ in range(k + m)]
        shards = [t[1:] for t in bufs]
        for j in range(k):
            shards[j].copy_(torch.from_numpy(data[j]))
        rs.encode_sep_route("bitsliced-fft", shards[:k], shards[k:])
        assert (to_host(shards[k:]) == want).all(), length
        assert rs.verify(shards), length
        shards[k + 7][length - 1] ^= 1
        assert not rs.verify(shards), length


@pytest.mark.parametrize("route", ["table", "mfma"])
@pytest.mark.parametrize("k,m", [(10, 4), (7, 13), (10, 10), (32, 32), (17, 5), (3, 2), (29, 9)])
def test_reconstruct_apply_routes(trs, k, m, route):
    """The fused table kernel and the int8-MFMA kernel (bit-expanded plan)
    rebuild every erased shard byte for byte, across tails, lengths that are
    not multiples of 4 and erasure counts 1..m, plus data-only."""
    rng = np.random.default_rng(k * 100 + m)
    G = O.build_matrix(8, k, k + m)
    rs = trs.ReedSolomon(k, m)
    rs.set_apply_route(route)
    for length in (1, 3, 4, 5, 127, 128, 129, 2048, 2049, 70003, (128 << 10) + 5):
        data = seeded(k, length, 99 + length)
        original = np.concatenate([data, O.rs8_encode(G, k, data)])
        e = int(rng.integers(1, m + 1))
        erased = set(rng.choice(k + m, e, replace=False).tolist())
        cur = rows_to_dev(original)
        rs.reconstruct(_erase_and_view(cur, erased))
        assert (to_host(cur) == original).all(), (k, m, length, sorted(erased), route)
        cur = rows_to_dev(original)
        rs.reconstruct_data(_erase_and_view(cur, erased))
        got = to_host(cur)
        for i in range(k + m):
            want = original[i] if (i < k or i not in erased) else np.zeros(length, np.uint8)
            assert (got[i] == want).all(), (k, m, length, i, route)


@pytest.mark.parametrize("k,m", [(5, 3), (29, 9)])
def test_auto_apply_route_by_length(trs, k, m):
    """Automatic routing of single-call matrix applies (shapes without a
    generated program, k <= 32): int8 MFMA below 512 KiB shards, the
    bit-sliced planes kernel above; encode and reconstruct match the oracle on
    both sides."""
    G = O.build_matrix(8, k, k + m)
    rs = trs.ReedSolomon(k, m)
    assert rs.encode_route(65536) == "gpu-int8-mfma-matrix"
    assert rs.encode_route(1 << 20) == "gpu-bitsliced-planes-matrix"
    rs.set_apply_route("table")
    assert rs.encode_route(65536) == "gpu-fused-matrix"
    rs.set_apply_route("auto")
    rng = np.random.default_rng(k)
    for length in (65536 + 3, (512 << 10) - 16, 512 << 10, (1 << 20) + 5):
        data = seeded(k, length, 7 + length)
        original = np.concatenate([data, O.rs8_encode(G, k, data)])
        cur = rows_to_dev(np.concatenate([data, np.zeros((m, length), np.uint8)]))
        rs.encode(list(cur))
        assert (to_host(cur) == original).all(), (k, m, length)
        erased = set(rng.choice(k + m, m, replace=False).tolist())
        rs.reconstruct(_erase_and_view(cur, erased))
        assert (to_host(cur) == original).all(), (k, m, length, sorted(erased))


def test_unaligned_rows_on_mfma_route(trs):
    # rows at odd offsets exercise the byte path of the MFMA kernel
    k, m, row_len = 10, 6, 1001
    rs = trs.ReedSolomon(k, m)
    rs.set_apply_route("mfma")
    G = O.build_matrix(8, k, k + m)
    data = seeded(k, row_len, 5)
    pristine = np.concatenate([data, O.rs8_encode(G, k, data)]).reshape(-1)
    rows = pristine.copy()
    present = [True] * (k + m)
    for e in (1, 2, 3, 11, 15):
        present[e] = False
        rows[e * row_len:(e + 1) * row_len] = 0
    drows = to_dev(rows)
    rs.reconstruct_rows(drows, row_len, present)
    assert (drows.cpu().numpy() == pristine).all()


@pytest.mark.parametrize("k,m", [(32, 32), (16, 16)])
def test_planes_tile_many_coefficients(trs, k, m):
    """The bit-sliced planes tile's 16-byte paths (nibble-wise
    multiply-accumulate, select-form transposes): single-call reconstructs of
    ~150 KiB shards (full 8 KiB tiles and a masked last tile) over 24 random
    erasure patterns, up to four 8-row output tiles each, rebuild every shard
    byte for byte.  The patterns' decode rows hold most of the 256 byte
    values, in every row position of a tile."""
    rng = np.random.default_rng(k * 1000 + m)
    G = O.build_matrix(8, k, k + m)
    rs = trs.ReedSolomon(k, m)
    rs.set_apply_route("table")  # >= 128 KiB: the planes kernel
    length = 19 * 8192 + 48
    data = seeded(k, length, 0xC0EF + k)
    original = np.concatenate([data, O.rs8_encode(G, k, data)])
    seen = set()
    for _ in range(24):
        e = int(rng.integers(1, m + 1))
        erased = sorted(rng.choice(k + m, e, replace=False).tolist())
        present = [i not in erased for i in range(k + m)]
        valid, mdata, drows, _ = O.rs8_decode_plan(G, k, m, present)
        seen.update(int(c) for c in np.asarray(drows).ravel())
        cur = rows_to_dev(original)
        rs.reconstruct(_erase_and_view(cur, set(erased)))
        assert (to_host(cur) == original).all(), (k, m, erased)
    assert len(seen) > 200, len(seen)
