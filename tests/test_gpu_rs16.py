"""GPU parity for ReedSolomon16 (GF((2^8)^2), split-plane shards) through the C
ABI, against the oracle: the frozen wire vector, the symbol-wise tower oracle
(src/reedsolomon16.rs:919-979), reconstruct patterns past 256 shards, verify,
prepared decoders and the even-length rule."""
import json
import os

import numpy as np
import pytest

import oracle as O
from util import deterministic_data, rows_to_dev, seeded, to_dev, to_host

pytestmark = pytest.mark.gpu

KATS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))


def encoded16(trs, k, m, length, seed=0x1234_5678):
    rs = trs.ReedSolomon16(k, m)
    data = deterministic_data(k * length, seed).reshape(k, length) if k * length <= 200_000 \
        else seeded(k, length, seed)
    shards = rows_to_dev(np.concatenate([data, np.zeros((m, length), np.uint8)]))
    rs.encode(shards)
    return rs, shards, data


def test_golden_wire_vector(trs):
    w = KATS["rs16_wire_vector"]
    k, m = w["k"], w["m"]
    rs = trs.ReedSolomon16(k, m)
    data = np.array(w["data"], np.uint8)
    shards = rows_to_dev(np.concatenate([data, np.zeros((m, data.shape[1]), np.uint8)]))
    rs.encode(shards)
    for idx, want in w["parity"].items():
        assert shards[int(idx)].cpu().numpy().tolist() == want


@pytest.mark.parametrize("k,m", [(4, 2), (7, 13), (10, 10), (17, 33)])
def test_encode_matches_symbolwise(trs, k, m):
    G = O.build_matrix(16, k, k + m)
    assert (trs.ReedSolomon16(k, m).generator() == G).all()
    for length in (2, 30, 32, 34, 62, 64, 66, 200, 2048, 65536):
        rs, shards, data = encoded16(trs, k, m, length, 0x9E37_79B9)
        want = O.rs16_symbolwise(G[k:], list(data), length)
        assert (to_host(shards[k:]) == want).all(), (k, m, length)


def test_encode_past_256_shards(trs):
    # coefficients outside the GF(2^8) subfield (rows with points >= 256)
    k, m = 200, 100
    G = O.build_matrix(16, k, k + m)
    for length in (34, 4096, 32768):
        data = seeded(k, length, 5)
        rs = trs.ReedSolomon16(k, m)
        shards = rows_to_dev(np.concatenate([data, np.zeros((m, length), np.uint8)]))
        rs.encode(shards)
        got = to_host(shards[k:])
        want = O.rs16_apply(G[k:], list(data), length)
        assert (got == want).all(), length


def test_encode_sep_and_matrix_route(trs):
    for k, m in [(4, 2), (17, 33)]:
        for length in (34, 512):
            rs, shards, data = encoded16(trs, k, m, length)
            parity = rows_to_dev(np.zeros((m, length), np.uint8))
            rs.encode_sep(rows_to_dev(data), parity)
            assert (to_host(parity) == to_host(shards[k:])).all()
            parity2 = rows_to_dev(np.zeros((m, length), np.uint8))
            rs.encode_sep_matrix(rows_to_dev(data), parity2)
            assert (to_host(parity2) == to_host(shards[k:])).all()


def _erase(shards, erased):
    view = []
    for i, s in enumerate(shards):
        if i in erased:
            s.zero_()
        view.append((s, i not in erased))
    return view


@pytest.mark.parametrize("k,m", [(4, 2), (17, 33), (10, 10), (100, 200)])
@pytest.mark.parametrize("length", [2, 34, 512])
def test_reconstruct_patterns(trs, k, m, length):
    # src/reedsolomon16.rs:1104-1140
    rs, shards, data = encoded16(trs, k, m, length)
    original = to_host(shards)
    total = k + m
    patterns = [[0], [k - 1], [k], list(range(min(m, k))), list(range(k, k + m)),
                [(e * 3) % total for e in range(m)]]
    for erased in patterns:
        for _ in range(2):
            cur = rows_to_dev(original)
            rs.reconstruct(_erase(cur, set(erased)))
            assert (to_host(cur) == original).all(), (k, m, length, erased)


def test_prepared_decoder_and_data_only(trs):
    k, m, length = 17, 33, 200
    rs, shards, data = encoded16(trs, k, m, length)
    original = to_host(shards)
    erased = {1, 3, 20, 40}
    dec = rs.prepare_decode([i not in erased for i in range(k + m)])
    for _ in range(3):
        cur = rows_to_dev(original)
        dec.reconstruct(_erase(cur, erased))
        assert (to_host(cur) == original).all()
    cur = rows_to_dev(original)
    with pytest.raises(trs.Error) as e:
        dec.reconstruct(_erase(cur, {0}))
    assert e.value.name == "InvalidShardFlags"

    k, m, length = 6, 4, 128
    rs, shards, data = encoded16(trs, k, m, length)
    original = to_host(shards)
    erased = {2, k + 1}
    dec = rs.prepare_decode([i not in erased for i in range(k + m)])
    cur = rows_to_dev(original)
    dec.reconstruct_data(_erase(cur, erased))
    got = to_host(cur)
    assert (got[2] == original[2]).all() and (got[k + 1] == 0).all()


def test_reconstruct_rows(trs):
    k, m, row_len = 10, 10, 258
    rs, shards, data = encoded16(trs, k, m, row_len)
    pristine = to_host(shards).reshape(-1)
    rows = pristine.copy()
    present = [True] * (k + m)
    for e in (0, 4, 12, 19):
        present[e] = False
        rows[e * row_len:(e + 1) * row_len] = 0
    drows = to_dev(rows)
    rs.reconstruct_rows(drows, row_len, present)
    assert (drows.cpu().numpy() == pristine).all()
    with pytest.raises(trs.Error) as e:
        rs.reconstruct_rows(drows, 257, present)
    assert e.value.name == "IncorrectShardSize"


def test_verify_blocks(trs):
    # src/reedsolomon16.rs:1197-1212
    k, m = 4, 3
    for length in (200, 2 * 16384, 2 * 16384 + 34):
        rs, shards, data = encoded16(trs, k, m, length)
        assert rs.verify(shards)
        host = to_host(shards)
        for pos in (0, length // 2, length - 1):
            bad = host.copy()
            bad[k, pos] ^= 1
            assert not rs.verify(rows_to_dev(bad)), (length, pos)


def test_odd_length_rejected(trs):
    rs = trs.ReedSolomon16(2, 2)
    shards = rows_to_dev(np.zeros((4, 5), np.uint8))
    with pytest.raises(trs.Error) as e:
        rs.encode(shards)
    assert e.value.name == "IncorrectShardSize"


def test_c3_shape_1024_1024(trs):
    """ReedSolomon16 (1024,1024) with 4 KiB split-plane shards (2048 symbols):
    the generator equals the oracle's, sampled parity rows equal the oracle's
    symbol-wise tower product, and erasing 1024 uniformly chosen shards per
    stripe reconstructs the original stripe."""
    import torch

    k = m = 1024
    L = 4096
    rs = trs.ReedSolomon16(k, m)
    G = O.build_matrix(16, k, k + m)
    assert (rs.generator() == G).all()
    rng = np.random.default_rng(0xC3)
    for stripe in range(2):
        buf = torch.empty((k + m) * L, dtype=torch.uint8, device="cuda")
        trs.fill_random(buf[:k * L], 0x7A9E5EED00000003 + stripe)
        shards = [buf[i * L:(i + 1) * L] for i in range(k + m)]
        rs.encode(shards)
        host = buf.cpu().numpy().reshape(k + m, L)
        rows = [0, 1, 255, 256, 511, 1023]
        want = O.rs16_symbolwise(G[[k + r for r in rows]], list(host[:k]), L)
        assert (host[[k + r for r in rows]] == want).all(), stripe
        original = buf.clone()
        erased = set(rng.choice(k + m, 1024, replace=False).tolist())
        view = []
        for i, s in enumerate(shards):
            if i in erased:
                s.zero_()
            view.append((s, i not in erased))
        rs.reconstruct(view)
        torch.cuda.synchronize()
        assert torch.equal(buf, original), stripe


def test_encode_sep_matrix_rejects_null_shard(trs):
    """A NULL data or parity pointer is an argument error, never a launch."""
    import ctypes

    k, m, L = 4, 2, 64
    rs = trs.ReedSolomon16(k, m)
    bufs = rows_to_dev(np.zeros((k + m, L), np.uint8))
    for hole in (1, k + 1):
        ptrs = [t.data_ptr() for t in bufs]
        ptrs[hole] = 0
        data = (ctypes.c_void_p * k)(*ptrs[:k])
        parity = (ctypes.c_void_p * m)(*ptrs[k:])
        dl = (ctypes.c_size_t * k)(*([L] * k))
        pl = (ctypes.c_size_t * m)(*([L] * m))
        rc = trs._lib.lib.trs_rs16_encode_sep_matrix(rs._h, data, dl, k, parity, pl, m, None)
        assert rc == -3, hole
