"""trs_rs16_encode_stripes_noalloc, the allocation-free entry point: its
output equals trs_rs16_encode_stripes' (tower-FFT route) byte for byte on two
calls in a row, the second smaller, and it rejects a request past its scratch
with TRS_ERR_SCRATCH_LIMIT; a small shape is also checked against the
oracle's symbol-wise parity."""
import ctypes

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

FFT_ROUTE = 2


def _table(buf):
    stripes, n, _ = buf.shape
    return (ctypes.c_void_p * (stripes * n))(
        *[buf[s, i].data_ptr() for s in range(stripes) for i in range(n)])


def test_noalloc_stripes_match_encode_stripes(trs):
    import torch

    lib = trs._lib.lib
    k, m = 1024, 1024
    rs = trs.ReedSolomon16(k, m)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for call, (stripes, length) in enumerate(((64, 4096), (16, 2048))):
        want = torch.empty((stripes, k + m, length), dtype=torch.uint8, device="cuda")
        trs.fill_random(want.reshape(-1), 0x5C + call)
        got = want.clone()
        got[:, k:].fill_(0xEE)  # stale bytes the call must overwrite
        assert lib.trs_rs16_encode_stripes(rs._h, _table(want), stripes, length, FFT_ROUTE,
                                           stream) == 0
        assert lib.trs_rs16_encode_stripes_noalloc(rs._h, _table(got), stripes, length,
                                                   stream) == 0
        torch.cuda.synchronize()
        assert torch.equal(got, want), (call, stripes, length)
    big = torch.zeros((1, k + m, 64), dtype=torch.uint8, device="cuda")
    one = [big[0, i].data_ptr() for i in range(k + m)]
    table = (ctypes.c_void_p * (65 * (k + m)))(*(one * 65))  # 65 stripes: past 1 MiB
    assert lib.trs_rs16_encode_stripes_noalloc(rs._h, table, 65, 64, stream) == -5
    assert "scratch" in trs._lib.lib.trs_status_message(-5).decode()


def test_noalloc_small_shape_matches_oracle(trs):
    import torch

    lib = trs._lib.lib
    k, m, stripes, length = 17, 33, 5, 600
    rs = trs.ReedSolomon16(k, m)
    G = O.build_matrix(16, k, k + m)
    data = [O.fill(900 + s, k * length).reshape(k, length) for s in range(stripes)]
    buf = torch.zeros((stripes, k + m, length), dtype=torch.uint8, device="cuda")
    for s in range(stripes):
        buf[s, :k] = torch.from_numpy(data[s]).cuda()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert lib.trs_rs16_encode_stripes_noalloc(rs._h, _table(buf), stripes, length, stream) == 0
    torch.cuda.synchronize()
    for s in range(stripes):
        want = O.rs16_symbolwise(G[k:], list(data[s]), length)
        assert (buf[s, k:].cpu().numpy() == want).all(), s
