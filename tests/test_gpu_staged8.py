"""GPU parity of the GF(2^8) staged FFT route (src/fft.rs:393-592 on the
byte-plane LDS interpreter) through the C ABI, against the oracle's generator-matrix encode.

Mirrors the reference's staged_matches_scalar test (src/reedsolomon.rs:997-1021:
staged and runtime shapes at ragged lengths) with the forced route, adds
full-size shards, unaligned shard pointers, the automatic route choice and the
reference-route report (src/reedsolomon.rs:239-253)."""
import numpy as np
import pytest

import oracle as O
from util import rows_to_dev, seeded, to_host

pytestmark = pytest.mark.gpu

# src/reedsolomon.rs:999-1006 plus the reference's staged shapes and wide ones
SHAPES = [(12, 8), (14, 14), (16, 16), (18, 6), (24, 8), (32, 32), (2, 2), (4, 4), (8, 8),
          (64, 64), (128, 32), (1, 4), (37, 21), (100, 156), (200, 55), (128, 128)]
LENGTHS = [1, 2, 3, 16, 31, 32, 33, 34, 100, 1000, 1001, 4096, 4098, 4099, 65536 + 64]


def _forced(trs, rs, k, m, data):
    length = data.shape[1]
    dev = rows_to_dev(np.concatenate([data, np.zeros((m, length), np.uint8)]))
    rs.encode_sep_route("staged-fft", dev[:k], dev[k:])
    return to_host(dev[k:])


@pytest.mark.parametrize("k,m", SHAPES)
def test_staged_route_matches_oracle(trs, k, m):
    print(f"staged route k={k} m={m}", flush=True)
    G = O.build_matrix(8, k, k + m)
    rs = trs.ReedSolomon(k, m)
    for length in LENGTHS:
        data = seeded(k, length, 0x51A6ED + length)
        assert (_forced(trs, rs, k, m, data) == O.rs8_encode(G, k, data)).all(), (k, m, length)


@pytest.mark.parametrize("k,m", [(64, 64), (8, 8), (128, 32)])
def test_staged_route_full_size(trs, k, m):
    # 1 MiB and 4 MiB + 32 shards: many workgroups, vector accesses
    G = O.build_matrix(8, k, k + m)
    rs = trs.ReedSolomon(k, m)
    for length in ((1 << 20), (4 << 20) + 32):
        if k * length > (256 << 20):
            continue
        data = seeded(k, length, 0xF00D + k)
        cols = slice(0, length, 4099)  # the oracle on a column sample keeps the check quick
        got = _forced(trs, rs, k, m, data)
        want = O.rs8_encode(G, k, np.ascontiguousarray(data[:, cols]))
        assert (got[:, cols] == want).all(), (k, m, length)
        # every column: the matrix route on the GPU is already pinned to the oracle
        dev = rows_to_dev(np.concatenate([data, np.zeros((m, length), np.uint8)]))
        rs.encode_sep_route("table", dev[:k], dev[k:])
        assert (to_host(dev[k:]) == got).all(), (k, m, length)


def test_staged_route_unaligned_pointers(trs):
    import torch

    k, m, length = 64, 64, 8192 + 64
    G = O.build_matrix(8, k, k + m)
    rs = trs.ReedSolomon(k, m)
    data = seeded(k, length, 0xA11)
    big = torch.zeros((k + m) * (length + 16), dtype=torch.uint8, device="cuda")
    views = [big[i * (length + 16) + 3: i * (length + 16) + 3 + length] for i in range(k + m)]
    for i in range(k):
        views[i].copy_(torch.from_numpy(data[i]))
    rs.encode_sep_route("staged-fft", views[:k], views[k:])
    assert (to_host(views[k:]) == O.rs8_encode(G, k, data)).all()


def test_staged_route_unaligned_odd_length(trs):
    import torch

    k, m, length = 40, 24, 12345
    G = O.build_matrix(8, k, k + m)
    rs = trs.ReedSolomon(k, m)
    data = seeded(k, length, 0xB11)
    big = torch.zeros((k + m) * (length + 16), dtype=torch.uint8, device="cuda")
    views = [big[i * (length + 16) + 5: i * (length + 16) + 5 + length] for i in range(k + m)]
    for i in range(k):
        views[i].copy_(torch.from_numpy(data[i]))
    rs.encode_sep_route("staged-fft", views[:k], views[k:])
    assert (to_host(views[k:]) == O.rs8_encode(G, k, data)).all()


def test_automatic_route_and_reports(trs):
    k, m = 64, 64
    G = O.build_matrix(8, k, k + m)
    rs = trs.ReedSolomon(k, m)
    assert rs.encode_route(1 << 20) == "gpu-staged-fft"
    assert rs.encode_route((1 << 20) + 1) == "gpu-staged-fft"
    assert rs.encode_route(4096) == "gpu-staged-fft"
    assert rs.reference_route(1 << 20) == "fft-staged"
    assert rs.reference_route(31) == "fused-matrix"
    for length in ((1 << 20), (1 << 20) + 1):
        data = seeded(k, length, 0xB0B + length)
        dev = rows_to_dev(np.concatenate([data, np.zeros((m, length), np.uint8)]))
        rs.encode(dev)
        cols = slice(0, length, 1021)
        want = O.rs8_encode(G, k, np.ascontiguousarray(data[:, cols]))
        assert (to_host(dev[k:])[:, cols] == want).all(), length
    # the device rule (k > 32): any length when the program's multiplies are
    # under 0.13 k*m (k = 64, 128) or 0.078 k*m (other k); else up to 1 MiB
    # (k = 64, 128) or 256 KiB; never for k <= 32
    r64_16 = trs.ReedSolomon(64, 16)
    assert r64_16.encode_route(1 << 20) == "gpu-staged-fft"
    assert r64_16.encode_route(4 << 20) != "gpu-staged-fft"
    assert trs.ReedSolomon(128, 128).encode_route(4 << 20) == "gpu-staged-fft"
    assert trs.ReedSolomon(100, 156).encode_route(4 << 20) == "gpu-staged-fft"
    r200 = trs.ReedSolomon(200, 55)
    assert r200.encode_route(256 << 10) == "gpu-staged-fft"
    assert r200.encode_route((256 << 10) + 1) != "gpu-staged-fft"
    assert r200.encode_route((256 << 10) - 1) == "gpu-staged-fft"
    assert trs.ReedSolomon(33, 16).encode_route(16 << 10) == "gpu-staged-fft"
    assert trs.ReedSolomon(8, 8).encode_route(1 << 20) != "gpu-staged-fft"
    assert trs.ReedSolomon(12, 8).encode_route(64 << 10) != "gpu-staged-fft"
    assert trs.ReedSolomon(31, 19).encode_route(64 << 10) != "gpu-staged-fft"
    assert trs.ReedSolomon(16, 16).reference_route(64) == "fft-generated"
    assert trs.ReedSolomon(12, 8).reference_route(1 << 20) == "fused-matrix"
    assert trs.ReedSolomon(128, 128).reference_route(1 << 20) == "fused-matrix"
    assert trs.ReedSolomon(128, 32).reference_route(1 << 20) == "fft-staged"


# ---- FFT erasure decoder over GF(2^8) (reconstruct route "fft") -----------

@pytest.mark.parametrize("k,m", [(4, 2), (10, 4), (7, 13), (32, 32), (64, 64), (100, 156),
                                 (200, 55), (128, 128)])
def test_fft_decode_route_matches_oracle(trs, k, m):
    print(f"fft decode k={k} m={m}", flush=True)
    n = k + m
    G = O.build_matrix(8, k, n)
    rs = trs.ReedSolomon(k, m)
    rs.set_decode_route("fft")
    rng = np.random.default_rng(k * 7 + m)
    for length in (1, 33, 4096, 65536 + 3):
        data = seeded(k, length, 0xDEC0 + length)
        full = np.concatenate([data, O.rs8_encode(G, k, data)])
        for data_only in (False, True):
            e = int(rng.integers(1, m + 1))
            erased = set(rng.choice(n, e, replace=False).tolist())
            shards = rows_to_dev(full)
            for i in erased:
                shards[i] = None
            (rs.reconstruct_data if data_only else rs.reconstruct)(shards)
            for i in range(n):
                if data_only and i >= k and i in erased:
                    assert shards[i] is None
                else:
                    assert (shards[i].cpu().numpy() == full[i]).all(), \
                        (k, m, length, sorted(erased), i)


def test_fft_decode_prepared_and_full_size(trs):
    k, m, length = 64, 64, (1 << 20) + 7
    n = k + m
    G = O.build_matrix(8, k, n)
    rs = trs.ReedSolomon(k, m)
    rs.set_decode_route("fft")
    data = seeded(k, length, 0xFEED)
    dev = rows_to_dev(np.concatenate([data, np.zeros((m, length), np.uint8)]))
    rs.encode(dev)
    ref = to_host(dev)
    cols = slice(0, length, 977)
    assert (ref[k:, cols] == O.rs8_encode(G, k, np.ascontiguousarray(data[:, cols]))).all()
    erased = set(range(0, n, 2))  # 64 erasures, data and parity
    present = [i not in erased for i in range(n)]
    dec = rs.prepare_decode(present)
    shards = list(dev)
    for i in erased:
        shards[i] = None
    dec.reconstruct(shards)
    assert (to_host(shards) == ref).all()
    # the matrix route gives the same bytes
    rs.set_decode_route("matrix")
    for i in erased:
        shards[i] = None
    rs.reconstruct(shards)
    assert (to_host(shards) == ref).all()


def test_automatic_decode_route_matches_oracle(trs):
    # k > 32 at 16 KiB / 200 KiB with >= 4 outputs: the automatic route takes
    # the FFT decoder; 1 MiB stays on the matrix plans -- every case byte-exact
    k, m = 48, 16
    n = k + m
    G = O.build_matrix(8, k, n)
    rs = trs.ReedSolomon(k, m)
    rng = np.random.default_rng(5)
    for length, e in ((16 << 10, 1), (16 << 10, 16), (200 << 10, 5), (1 << 20, 9)):
        data = seeded(k, length, 0xA070 + length + e)
        full = np.concatenate([data, O.rs8_encode(G, k, data)])
        erased = set(rng.choice(n, e, replace=False).tolist())
        shards = rows_to_dev(full)
        for i in erased:
            shards[i] = None
        rs.reconstruct(shards)
        assert (to_host(shards) == full).all(), (length, sorted(erased))
