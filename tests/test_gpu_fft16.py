"""GPU parity of the tower-FFT route (src/fft16.rs:887-930 is the reference
path): the interpreter kernel's parity equals the symbol-wise oracle on the
reference's staged_matches_matrix shapes and more, across plane lengths that
end mid-group, with unaligned shards, in multi-stripe launches, and at the
(1024,1024) 4 KiB shape against the matrix route."""
import numpy as np
import pytest

import oracle as O
from util import rows_to_dev, seeded, to_host

pytestmark = pytest.mark.gpu

SHAPES = [(2, 2), (3, 5), (5, 3), (7, 13), (16, 16), (17, 33), (32, 96), (48, 144), (63, 65),
          (86, 170), (100, 50), (128, 128), (4, 2), (10, 4), (11, 20), (33, 31), (255, 1),
          (2, 1), (300, 212), (513, 511), (1000, 24)]


def _fft_parity(trs, rs, data, m):
    length = data.shape[1]
    parity = rows_to_dev(np.zeros((m, length), np.uint8))
    rs.encode_sep_fft(rows_to_dev(data), parity)
    return to_host(parity)


@pytest.mark.parametrize("k,m", SHAPES)
def test_fft_route_matches_symbolwise(trs, k, m):
    G = O.build_matrix(16, k, k + m)
    rs = trs.ReedSolomon16(k, m)
    for length in (2, 66, 4096, 6150):
        print(f"fft16 k={k} m={m} len={length}", flush=True)
        data = seeded(k, length, k * 131 + length)
        want = O.rs16_symbolwise(G[k:], list(data), length)
        assert (_fft_parity(trs, rs, data, m) == want).all(), (k, m, length)


def test_fft_route_unaligned_shards(trs):
    import torch

    k, m, length = 17, 33, 2 * 100
    G = O.build_matrix(16, k, k + m)
    rs = trs.ReedSolomon16(k, m)
    data = seeded(k, length, 77)
    buf = torch.zeros((k + m) * (length + 3) + 1, dtype=torch.uint8, device="cuda")
    shards = [buf[1 + i * (length + 3):1 + i * (length + 3) + length] for i in range(k + m)]
    for i in range(k):
        shards[i].copy_(torch.from_numpy(data[i]).cuda())
    rs.encode_sep_fft(shards[:k], shards[k:])
    want = O.rs16_symbolwise(G[k:], list(data), length)
    assert (to_host(shards[k:]) == want).all()


def test_routes(trs):
    # GPU routing: the FFT whenever the program fits (the reference's x86
    # multiply-count rule would keep (11,20) on the matrix route).
    assert trs.ReedSolomon16(11, 20).encode_route() == "fft"
    assert trs.ReedSolomon16(17, 33).encode_route() == "fft"
    assert trs.ReedSolomon16(1024, 1024).encode_route() == "fft"
    assert trs.ReedSolomon16(1, 5).encode_route() == "matrix"
    rs = trs.ReedSolomon16(8, 4100)  # no program past k + m = 4096
    assert rs.encode_route() == "matrix"
    data = rows_to_dev(np.zeros((8, 4), np.uint8))
    parity = rows_to_dev(np.zeros((4100, 4), np.uint8))
    with pytest.raises(trs.Error) as e:
        rs.encode_sep_fft(data, parity)
    assert e.value.code == -4


def test_encode_stripes_routes_agree(trs):
    import torch

    k, m, length, stripes = 86, 170, 4096, 5
    G = O.build_matrix(16, k, k + m)
    rs = trs.ReedSolomon16(k, m)
    bufs = {}
    for route in ("auto", "matrix", "fft"):
        buf = torch.zeros((stripes, k + m, length), dtype=torch.uint8, device="cuda")
        for s in range(stripes):
            trs.fill_random(buf[s, :k].reshape(-1), 0xF0 + s)
        rs.encode_stripes([[buf[s, i] for i in range(k + m)] for s in range(stripes)], route)
        bufs[route] = buf.cpu().numpy()
    assert (bufs["auto"] == bufs["fft"]).all()
    assert (bufs["matrix"] == bufs["fft"]).all()
    for s in (0, stripes - 1):
        want = O.rs16_symbolwise(G[k:], list(bufs["fft"][s, :k]), length)
        assert (bufs["fft"][s, k:] == want).all(), s


def test_c3_workload_fft_equals_matrix(trs):
    """(1024,1024), 4 KiB shards, 64 stripes in one FFT launch and through the
    matrix route: byte-identical; sampled columns of two stripes equal the
    symbol-wise oracle."""
    import torch

    k = m = 1024
    L, stripes = 4096, 64
    rs = trs.ReedSolomon16(k, m)
    buf = torch.zeros((stripes, k + m, L), dtype=torch.uint8, device="cuda")
    trs.fill_random(buf.reshape(-1), 0x7A9E5EED00000003)
    ref = buf.clone()
    views = [[buf[s, i] for i in range(k + m)] for s in range(stripes)]
    rs.encode_stripes(views, "fft")
    fft = buf[:, k:].clone()
    buf.copy_(ref)
    rs.encode_stripes(views, "matrix")
    torch.cuda.synchronize()
    assert torch.equal(fft, buf[:, k:])
    G = O.build_matrix(16, k, k + m)
    cols = np.r_[0:40, 2000:2048]  # symbol columns, both plane halves
    host = buf.cpu().numpy()
    for s in (0, stripes - 1):
        sub = np.ascontiguousarray(
            np.concatenate([host[s, :k][:, cols], host[s, :k][:, L // 2 + cols]], axis=1))
        want = O.rs16_symbolwise(G[k:], list(sub), sub.shape[1])
        got = np.concatenate([host[s, k:][:, cols], host[s, k:][:, L // 2 + cols]], axis=1)
        assert (got == want).all(), s


@pytest.mark.parametrize("k,m", [(2, 2), (4, 2), (10, 4), (7, 13), (17, 33), (32, 32), (100, 50),
                                 (200, 56), (300, 212), (1000, 24)])
def test_fft_decode_route_restores_stripe(trs, k, m):
    """The additive-FFT erasure decoder and the matrix route rebuild the same
    (original) shards; data-only leaves erased parity untouched."""
    import torch

    n = k + m
    rng = np.random.default_rng(k * 13 + m)
    for length in (2, 130, 4096):
        print(f"fft16 decode k={k} m={m} len={length}", flush=True)
        rs = trs.ReedSolomon16(k, m)
        buf = torch.zeros((n, length), dtype=torch.uint8, device="cuda")
        trs.fill_random(buf[:k].reshape(-1), k * 1000 + length)
        rs.encode([buf[i] for i in range(n)])
        pristine = buf.clone()
        for e in sorted({1, m, int(rng.integers(1, m + 1))}):
            erased = rng.choice(n, e, replace=False).tolist()
            for route in ("fft", "matrix"):
                rs.set_decode_route(route)
                buf[erased] = 0
                rs.reconstruct([(buf[i], i not in erased) for i in range(n)])
                torch.cuda.synchronize()
                assert torch.equal(buf, pristine), (route, length, sorted(erased))
            rs.set_decode_route("fft")
            buf[erased] = 0
            rs.reconstruct_data([(buf[i], i not in erased) for i in range(n)])
            torch.cuda.synchronize()
            for i in range(n):
                if i < k or i not in erased:
                    assert torch.equal(buf[i], pristine[i]), (i, sorted(erased))
                else:
                    assert not buf[i].any()
            buf.copy_(pristine)


def test_fft_decode_prepared_and_rows(trs):
    import torch

    k, m, length = 200, 100, 512
    n = k + m
    rs = trs.ReedSolomon16(k, m)
    rs.set_decode_route("fft")
    rows = torch.zeros(n * length, dtype=torch.uint8, device="cuda")
    trs.fill_random(rows[:k * length], 99)
    shards = [rows[i * length:(i + 1) * length] for i in range(n)]
    rs.encode(shards)
    pristine = rows.clone()
    erased = list(range(0, 300, 3))
    present = [i not in erased for i in range(n)]
    dec = rs.prepare_decode(present)
    for i in erased:
        shards[i].zero_()
    dec.reconstruct([(shards[i], present[i]) for i in range(n)])
    torch.cuda.synchronize()
    assert torch.equal(rows, pristine)
    for i in erased:
        shards[i].zero_()
    rs.reconstruct_rows(rows, length, present)
    torch.cuda.synchronize()
    assert torch.equal(rows, pristine)


def test_c3_reconstruct_fft_route(trs):
    """(1024,1024), 4 KiB shards: 1024 uniformly chosen erasures per stripe
    rebuilt by the FFT decoder equal the original stripe (two patterns)."""
    import torch

    k = m = 1024
    L = 4096
    rs = trs.ReedSolomon16(k, m)
    rs.set_decode_route("fft")
    buf = torch.zeros((k + m, L), dtype=torch.uint8, device="cuda")
    trs.fill_random(buf[:k].reshape(-1), 0xC3C3)
    rs.encode([buf[i] for i in range(k + m)])
    pristine = buf.clone()
    rng = np.random.default_rng(0xC3)
    for _ in range(2):
        erased = rng.choice(k + m, 1024, replace=False).tolist()
        buf[erased] = 0
        rs.reconstruct([(buf[i], i not in erased) for i in range(k + m)])
        torch.cuda.synchronize()
        assert torch.equal(buf, pristine)


@pytest.mark.parametrize("k,m,stripes", [(17, 33, 6), (200, 100, 5), (1024, 1024, 4)])
def test_reconstruct_stripes_each_pattern(trs, k, m, stripes):
    """Many stripes with their own erasure patterns in one call (FFT stripes
    share one launch) equal per-stripe reconstructs and the original data;
    data-only leaves erased parity untouched."""
    import torch

    n, L = k + m, 4096
    rs = trs.ReedSolomon16(k, m)
    buf = torch.zeros((stripes, n, L), dtype=torch.uint8, device="cuda")
    for s in range(stripes):
        trs.fill_random(buf[s, :k].reshape(-1), 0x51 + s)
    rs.encode_stripes(buf)
    pristine = buf.clone()
    rng = np.random.default_rng(k + m)
    present = np.ones((stripes, n), bool)
    for s in range(stripes):
        present[s, rng.choice(n, int(rng.integers(1, m + 1)), replace=False)] = False
    present[0] = True  # a complete stripe is a no-op
    for route in ("fft", "matrix", "auto"):
        rs.set_decode_route(route)
        buf[torch.from_numpy(~present).cuda()] = 0
        rs.reconstruct_stripes(buf, present)
        torch.cuda.synchronize()
        assert torch.equal(buf, pristine), route
    rs.set_decode_route("fft")
    buf[torch.from_numpy(~present).cuda()] = 0
    rs.reconstruct_stripes(buf, present, data_only=True)
    torch.cuda.synchronize()
    missing_parity = torch.from_numpy(~present).cuda()
    missing_parity[:, :k] = False
    assert not buf[missing_parity].any()
    keep = ~missing_parity
    assert torch.equal(buf[keep], pristine[keep])


def test_reconstruct_stripes_irregular_pointers(trs):
    """Shards that are not one (stripes, n, len) tensor -- here every shard a
    separate allocation, and one stripe's shards in reverse memory order --
    take the per-call pointer table instead of the plan's indexed addressing;
    both layouts in one call rebuild every erased shard."""
    import ctypes

    import torch

    k, m, L, stripes = 17, 33, 4096, 3
    n = k + m
    rs = trs.ReedSolomon16(k, m)
    G = O.build_matrix(16, k, n)
    data = [seeded(k, L, 0x77 + s) for s in range(stripes)]
    full = [np.concatenate([d, O.rs16_symbolwise(G[k:], list(d), L)]) for d in data]
    shards = [[torch.from_numpy(full[s][i].copy()).cuda() for i in range(n)]
              for s in range(stripes)]
    rev = torch.from_numpy(np.ascontiguousarray(full[2][::-1])).cuda()  # stripe 2: reversed rows
    shards[2] = [rev[n - 1 - i] for i in range(n)]
    rng = np.random.default_rng(5)
    present = np.ones((stripes, n), np.uint8)
    for s in range(stripes):
        present[s, rng.choice(n, m, replace=False)] = 0
    for s in range(stripes):
        for i in range(n):
            if not present[s, i]:
                shards[s][i].zero_()
    ptrs = (ctypes.c_void_p * (stripes * n))(*[t.data_ptr() for st in shards for t in st])
    rs.set_decode_route("fft")
    rc = trs._lib.lib.trs_rs16_reconstruct_stripes(
        rs._h, ptrs, present.ctypes.data_as(trs._lib.c_u8p), stripes, L, 0,
        ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    torch.cuda.synchronize()
    for s in range(stripes):
        got = np.stack([t.cpu().numpy() for t in shards[s]])
        assert (got == full[s]).all(), s
