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
    assert trs.ReedSolomon16(11, 20).encode_route() == "matrix"
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
