"""verify on device shards through every encode route (src/reedsolomon.rs:430-463,
src/reedsolomon16.rs:453-505): the generated bit-sliced programs and the
staged / tower FFT programs compare their outputs with the parity shards in
their store epilogue, the matrix route uses the apply kernel's verify mode.
Each must answer true on a valid stripe and false after a single-bit flip at
the start, middle or end of any parity shard (the reference's verify test,
src/reedsolomon.rs:1188-1207)."""
import numpy as np
import pytest

import oracle as O
from util import rows_to_dev, seeded

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k,m,route", [
    (7, 13, "gpu-bitsliced-fft"), (32, 32, "gpu-bitsliced-fft"),
    (10, 4, "gpu-bitsliced-matrix"), (48, 16, "gpu-staged-fft"),
    (64, 64, "gpu-staged-fft"), (20, 10, None)])
def test_verify_every_route(trs, k, m, route):
    rs = trs.ReedSolomon(k, m)
    G = O.build_matrix(8, k, k + m)
    checked = 0
    for length in (1, 100, 8192, 8192 + 17, 70_000, 256 * 1024):
        if route is not None and rs.encode_route(length) != route:
            continue
        checked += 1
        data = seeded(k, length, 400 + length)
        full = np.concatenate([data, O.rs8_encode(G, k, data)])
        assert rs.verify(rows_to_dev(full)), (k, m, length)
        for j, pos in ((0, 0), (m // 2, length // 2), (m - 1, length - 1)):
            bad = full.copy()
            bad[k + j, pos] ^= 1
            assert not rs.verify(rows_to_dev(bad)), (length, j, pos)
    assert checked >= 3, (k, m, route)


def test_verify_rs16_fft_route_c3_shape(trs):
    """(1024,1024) on the tower FFT route: flips in the low and the high plane
    of the first, a middle and the last parity shard."""
    import torch

    k, m, length = 1024, 1024, 4096
    rs = trs.ReedSolomon16(k, m)
    assert rs.encode_route() == "fft"
    buf = torch.empty((k + m, length), dtype=torch.uint8, device="cuda")
    trs.fill_random(buf[:k].reshape(-1), 0xC3)
    shards = [buf[i] for i in range(k + m)]
    rs.encode(shards)
    torch.cuda.synchronize()
    assert rs.verify(shards)
    for j, pos in ((k, 0), (k + 511, 2047), (k + 511, 2048), (k + m - 1, length - 1)):
        buf[j, pos] ^= 4
        assert not rs.verify(shards), (j, pos)
        buf[j, pos] ^= 4
    assert rs.verify(shards)
    # a shape without a staged program (k > 1024) verifies on the matrix route
    k2, m2, l2 = 1030, 6, 64
    rs_m = trs.ReedSolomon16(k2, m2)
    assert rs_m.encode_route() == "matrix"
    data = seeded(k2, l2, 5)
    G = O.build_matrix(16, k2, k2 + m2)
    full = np.concatenate([data, O.rs16_symbolwise(G[k2:], list(data), l2)])
    assert rs_m.verify(rows_to_dev(full))
    full[k2 + 3, 40] ^= 2
    assert not rs_m.verify(rows_to_dev(full))
