"""The tower on int8 MFMA: a split-plane GF((2^8)^2) shard is two byte shards
(its low and high planes), and multiplying by a tower constant is GF(2)-linear
on a symbol's 16 bits, so an m x k tower matrix runs as the 16m x 16k bit
matrix of the byte problem on v_mfma_i32_32x32x32_i8, keeping the low bit of
each exact int32 sum.  Encode (encode_sep_matrix and the matrix route),
reconstruct on the matrix plans (codec and prepared decoders, full and
data-only) and the batched stripe calls must equal the symbol-wise tower
oracle (src/reedsolomon16.rs:919-979) byte for byte, at plane lengths that are
and are not multiples of 4 (the kernel's aligned and byte paths) and across
its 128-column blocks and 2048-column workgroup tiles."""
import numpy as np
import pytest

import oracle as O
from util import rows_to_dev, seeded, to_host

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1), (2, 1), (3, 2), (4, 4), (10, 4), (7, 9), (16, 4), (16, 16), (13, 40)]
LENGTHS = [2, 6, 64, 4098, 70_000, 1 << 20]


@pytest.mark.parametrize("k,m", SHAPES)
def test_mfma_encode_matches_oracle(trs, k, m):
    rs = trs.ReedSolomon16(k, m)
    rs.set_apply_route("mfma")
    G = O.build_matrix(16, k, k + m)
    for length in LENGTHS:
        data = seeded(k, length, 0x16AF + length)
        want = O.rs16_symbolwise(G[k:], list(data), length)
        d = rows_to_dev(data)
        p = rows_to_dev(np.zeros((m, length), np.uint8))
        rs.encode_sep_matrix(d, p)
        assert (to_host(p) == want).all(), (k, m, length)
        # the table kernel writes the same bytes
        rs.set_apply_route("table")
        q = rows_to_dev(np.zeros((m, length), np.uint8))
        rs.encode_sep_matrix(d, q)
        assert (to_host(q) == want).all(), (k, m, length)
        rs.set_apply_route("mfma")


@pytest.mark.parametrize("k,m", [(2, 1), (4, 4), (10, 4), (16, 16), (13, 40)])
def test_mfma_reconstruct_matches_oracle(trs, k, m):
    rs = trs.ReedSolomon16(k, m)
    rs.set_decode_route("matrix")
    rs.set_apply_route("mfma")
    G = O.build_matrix(16, k, k + m)
    rng = np.random.default_rng(k * 131 + m)
    for length in (6, 4098, 70_000, 1 << 20):
        data = seeded(k, length, 77 + length)
        full = np.concatenate([data, O.rs16_symbolwise(G[k:], list(data), length)])
        for e in sorted({1, m // 2 or 1, m}):
            erased = sorted(rng.choice(k + m, e, replace=False).tolist())
            present = [i not in erased for i in range(k + m)]
            start = full.copy()
            start[erased] = 0
            sh = rows_to_dev(start)
            rs.reconstruct([(t, p) for t, p in zip(sh, present)])
            assert (to_host(sh) == full).all(), (k, m, length, erased)
            # data only: missing parity stays as it was
            sh = rows_to_dev(start)
            rs.reconstruct_data([(t, p) for t, p in zip(sh, present)])
            got = to_host(sh)
            assert (got[:k] == full[:k]).all(), (k, m, length, erased)
            assert (got[k:][[not p for p in present[k:]]] == 0).all()
            # a decoder prepared now keeps the MFMA route
            dec = rs.prepare_decode(present)
            sh = rows_to_dev(start)
            dec.reconstruct([(t, p) for t, p in zip(sh, present)])
            assert (to_host(sh) == full).all(), (k, m, length, erased)


def test_mfma_stripe_calls_and_limits(trs):
    """The multi-stripe encode and reconstruct on the matrix route use the
    apply route too; shapes beyond the kernel (k > 16, or more than 256 byte
    shards) refuse the forced route and keep the table kernel."""
    import torch

    k, m, S, length = 10, 6, 5, 8192
    rs = trs.ReedSolomon16(k, m)
    rs.set_apply_route("mfma")
    rs.set_decode_route("matrix")
    G = O.build_matrix(16, k, k + m)
    buf = torch.zeros((S, k + m, length), dtype=torch.uint8, device="cuda")
    data = [seeded(k, length, 900 + s) for s in range(S)]
    for s in range(S):
        buf[s, :k] = torch.from_numpy(data[s]).cuda()
    rs.encode_stripes(buf, "matrix")
    torch.cuda.synchronize()
    full = [np.concatenate([d, O.rs16_symbolwise(G[k:], list(d), length)]) for d in data]
    for s in range(S):
        assert (buf[s].cpu().numpy() == full[s]).all(), s
    rng = np.random.default_rng(5)
    present = np.ones((S, k + m), bool)
    for s in range(S):
        present[s, rng.choice(k + m, 1 + s, replace=False)] = False
    for s in range(S):
        buf[s][torch.from_numpy(~present[s]).cuda()] = 0
    rs.reconstruct_stripes(buf, present)
    torch.cuda.synchronize()
    for s in range(S):
        assert (buf[s].cpu().numpy() == full[s]).all(), s

    for shape in ((17, 4), (16, 113)):
        with pytest.raises(trs.Error):
            trs.ReedSolomon16(*shape).set_apply_route("mfma")


def test_automatic_routes_take_matrix_kernels_on_short_data(trs):
    """With no route forced, a short-data tower encode runs MFMA from 512 KiB
    (m <= 4) and the Karatsuba kernel from 4 MiB, and reconstructs on matrix
    plans run MFMA on short shards: every result equals the oracle."""
    for (k, m), length in (((10, 4), 1 << 20), ((10, 4), (4 << 20) + 2), ((4, 4), 4 << 20),
                           ((16, 2), 600_000)):
        rs = trs.ReedSolomon16(k, m)
        G = O.build_matrix(16, k, k + m)
        data = seeded(k, length, 0xA070 + k)
        full = np.concatenate([data, O.rs16_symbolwise(G[k:], list(data), length)])
        sh = rows_to_dev(np.concatenate([data, np.zeros((m, length), np.uint8)]))
        rs.encode(sh)
        assert (to_host(sh) == full).all(), (k, m, length)
        assert rs.verify(sh)
        erased = [0, k]  # one data and one parity shard
        present = [i not in erased for i in range(k + m)]
        start = full.copy()
        start[erased] = 0
        sh = rows_to_dev(start)
        rs.reconstruct([(t, p) for t, p in zip(sh, present)])
        assert (to_host(sh) == full).all(), (k, m, length)
