"""Host-resident encode / reconstruct (pipelined through pinned staging), the
Option<T> allocate-missing reconstruct and the u16 <-> split-plane wire
conversions, against the oracle."""
import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu


def _stripe(k, m, length, seed):
    G = O.build_matrix(8, k, k + m)
    data = O.fill(seed, k * length).reshape(k, length)
    return G, data, O.rs8_encode(G, k, data)


# (48, 16): k > 32, so chunks up to 256 KiB take the staged FFT route
@pytest.mark.parametrize("k,m", [(10, 4), (7, 13), (20, 10), (48, 16)])
@pytest.mark.parametrize("pinned", [False, True])
def test_encode_host(trs, k, m, pinned):
    try:
        for length, chunk in ((1, 0), (1000, 0), (50_000, 4096), (50_001, 4096),
                              (300_000, 65536 + 64)):
            trs.set_host_chunk(chunk)
            G, data, want = _stripe(k, m, length, 11 + length)
            if pinned:
                d = [torch.from_numpy(data[i].copy()).pin_memory() for i in range(k)]
                p = [torch.zeros(length, dtype=torch.uint8).pin_memory() for _ in range(m)]
                got = lambda: np.stack([t.numpy() for t in p])
            else:
                d = [data[i].copy() for i in range(k)]
                p = [np.zeros(length, np.uint8) for _ in range(m)]
                got = lambda: np.stack(p)
            trs.ReedSolomon(k, m).encode_sep(d, p)
            assert (got() == want).all(), (k, m, length, chunk, pinned)
    finally:
        trs.set_host_chunk(0)


@pytest.mark.parametrize("k,m", [(10, 4), (20, 10)])
def test_reconstruct_host(trs, k, m):
    rng = np.random.default_rng(k)
    rs = trs.ReedSolomon(k, m)
    try:
        for length, chunk in ((777, 0), (100_000, 8192)):
            trs.set_host_chunk(chunk)
            G, data, parity = _stripe(k, m, length, 5 + length)
            full = np.concatenate([data, parity])
            for data_only in (False, True):
                erased = set(rng.choice(k + m, m, replace=False).tolist())
                rows = [full[i].copy() for i in range(k + m)]
                for i in erased:
                    rows[i][:] = 0
                present = [i not in erased for i in range(k + m)]
                shards = [(r, p) for r, p in zip(rows, present)]
                (rs.reconstruct_data if data_only else rs.reconstruct)(shards)
                for i in range(k + m):
                    if data_only and i >= k and i in erased:
                        assert not rows[i].any()
                    else:
                        assert (rows[i] == full[i]).all(), (length, sorted(erased), i)
    finally:
        trs.set_host_chunk(0)
    rows = [np.zeros(64, np.uint8) for _ in range(k + m)]
    with pytest.raises(trs.Error) as e:
        rs.reconstruct([(r, i < k - 1) for i, r in enumerate(rows)])
    assert e.value.name == "TooFewShardsPresent"


def test_reconstruct_alloc(trs):
    k, m, length = 10, 4, 70_001
    G, data, parity = _stripe(k, m, length, 99)
    full = np.concatenate([data, parity])
    rs = trs.ReedSolomon(k, m)
    for data_only in (False, True):
        shards = [torch.from_numpy(full[i].copy()).cuda() for i in range(k + m)]
        missing = {1, 7, 11, 13}
        for i in missing:
            shards[i] = None
        out = rs.reconstruct_alloc(shards, data_only)
        for i in range(k + m):
            if i in missing and data_only and i >= k:
                assert out[i] is None
            elif i in missing:
                assert isinstance(out[i], trs.DeviceBuffer) and out[i].nbytes == length
                assert (out[i].to_tensor().cpu().numpy() == full[i]).all(), (i, data_only)
                out[i].free()
            else:
                assert out[i] is shards[i]
    # errors leave nothing allocated and the entries None
    shards = [None] * (k + m)
    shards[0] = torch.zeros(16, dtype=torch.uint8, device="cuda")
    with pytest.raises(trs.Error) as e:
        rs.reconstruct_alloc(shards)
    assert e.value.name == "TooFewShardsPresent"


@pytest.mark.parametrize("n", [1, 7, 8, 9, 1000, 65536 + 3])
@pytest.mark.parametrize("offset", [0, 2])
def test_u16_split_round_trip(trs, n, offset):
    rng = np.random.default_rng(n + offset)
    raw = rng.integers(0, 256, 2 * n + offset, dtype=np.uint8)
    inter = torch.from_numpy(raw).cuda()[offset:]
    split = torch.zeros(2 * n + offset, dtype=torch.uint8, device="cuda")[offset:]
    trs.u16_to_split(inter, split)
    host = raw[offset:]
    want = np.concatenate([host[0::2], host[1::2]])
    assert (split.cpu().numpy() == want).all()
    back = torch.zeros(2 * n + offset, dtype=torch.uint8, device="cuda")[offset:]
    trs.split_to_u16(split, back)
    assert (back.cpu().numpy() == host).all()
