"""The batched GF(2^8) work list: every request's output in one launch is
byte-identical to the same request run alone, across mixed shapes, lengths
(including unaligned tails) and erasure patterns."""
import numpy as np
import pytest

import oracle as O
from util import rows_to_dev, seeded, to_host

pytestmark = pytest.mark.gpu

SHAPES = [(7, 13), (10, 10), (14, 14), (16, 16), (18, 6), (32, 32), (10, 4)]


@pytest.mark.parametrize("route", ["table", "mfma", "auto"])
def test_batch_matches_per_call(trs, route):
    import torch

    rng = np.random.default_rng(11)
    codecs = {s: trs.ReedSolomon(*s) for s in SHAPES}
    for c in codecs.values():
        c.set_apply_route(route)
    batch = trs.Batch()
    cases = []
    queued = 0
    for n in range(60):
        k, m = SHAPES[n % len(SHAPES)]
        length = int(rng.choice([1, 63, 64, 4096, 4097, 5000, 70000]))
        data = seeded(k, length, 1000 + n)
        G = O.build_matrix(8, k, k + m)
        full = np.concatenate([data, O.rs8_encode(G, k, data)])
        if n % 3 == 0:   # encode request
            shards = rows_to_dev(np.concatenate([data, np.zeros((m, length), np.uint8)]))
            batch.add_encode(codecs[(k, m)], shards)
            cases.append((shards, full, None))
            queued += 1
        else:            # reconstruct request
            e = int(rng.integers(1, m + 1))
            erased = set(rng.choice(k + m, e, replace=False).tolist())
            cur = full.copy()
            for i in erased:
                cur[i] = 0
            shards = rows_to_dev(cur)
            present = [i not in erased for i in range(k + m)]
            data_only = n % 7 == 0
            batch.add_reconstruct(codecs[(k, m)], shards, present, data_only)
            # a data-only request with no missing data has nothing to queue
            queued += 0 if (data_only and all(i >= k for i in erased)) else 1
            want = full.copy()
            if data_only:
                for i in erased:
                    if i >= k:
                        want[i] = 0
            cases.append((shards, want, erased))
    batch.prepare()
    batch.run()
    torch.cuda.synchronize()
    for shards, want, _ in cases:
        assert (to_host(shards) == want).all()
    assert len(batch) == queued
    assert batch.bytes() > 0


def test_batch_rerun_is_idempotent(trs):
    import torch

    rs = trs.ReedSolomon(10, 4)
    k, m, L = 10, 4, 8192
    data = seeded(k, L, 3)
    G = O.build_matrix(8, k, k + m)
    shards = rows_to_dev(np.concatenate([data, np.zeros((m, L), np.uint8)]))
    b = trs.Batch()
    b.add_encode(rs, shards)
    for _ in range(3):
        b.run()
    torch.cuda.synchronize()
    assert (to_host(shards[k:]) == O.rs8_encode(G, k, data)).all()


@pytest.mark.parametrize("k,m", [(32, 32), (16, 16)])
def test_batch_pair_program_tiles(trs, k, m):
    """The (32,32) batched encode runs on wave pairs, one per 2 KiB window of
    each 8 KiB tile, that swap registers through LDS between their halves of
    the program ((16,16), one program per wave, on the same cases).  Lengths
    end in each of the four windows, on full tiles, on aligned tails that are
    and are not multiples of 16, and shards start one byte past an
    allocation (byte accesses); all in one launch, against the oracle's
    generator-matrix encode."""
    import torch

    rs = trs.ReedSolomon(k, m)
    G = O.build_matrix(8, k, k + m)
    lengths = [1, 16, 2048, 2064, 4096 + 16, 6144 + 1, 8192, 8192 + 2048 + 16,
               3 * 8192 + 5000, 70000, 1 << 18]
    b = trs.Batch()
    cases = []
    for i, L in enumerate(lengths):
        for unaligned in (False, True):
            data = seeded(k, L, 500 + 2 * i + int(unaligned))
            if unaligned:
                bufs = [torch.full((L + 1,), 0x5A, dtype=torch.uint8, device="cuda")
                        for _ in range(k + m)]
                shards = [t[1:] for t in bufs]
                for j in range(k):
                    shards[j].copy_(torch.from_numpy(data[j]))
            else:
                shards = rows_to_dev(np.concatenate([data, np.full((m, L), 0x5A, np.uint8)]))
            b.add_encode(rs, shards)
            cases.append((L, unaligned, shards, O.rs8_encode(G, k, data)))
    b.prepare()
    b.run()
    torch.cuda.synchronize()
    for L, unaligned, shards, want in cases:
        got = to_host(shards[k:])
        assert (got == want).all(), (L, unaligned, np.argwhere(got != want)[:4].tolist())


def test_batch_graph_replays_match_eager(trs):
    """Graph mode: every replay of the captured launches (with eager work and
    synchronisations between replays) rebuilds the same bytes as the eager
    run, for a mixed encode + reconstruct batch on all three kernels."""
    import torch

    rng = np.random.default_rng(21)
    codecs = {s: trs.ReedSolomon(*s) for s in SHAPES}
    codecs[(10, 4)].set_apply_route("mfma")
    batch = trs.Batch()
    cases = []
    for n in range(28):
        k, m = SHAPES[n % len(SHAPES)]
        length = int(rng.choice([64, 4096, 5000, 70000]))
        data = seeded(k, length, 500 + n)
        full = np.concatenate([data, O.rs8_encode(O.build_matrix(8, k, k + m), k, data)])
        if n % 2 == 0:
            shards = rows_to_dev(np.concatenate([data, np.zeros((m, length), np.uint8)]))
            batch.add_encode(codecs[(k, m)], shards)
            wipe = list(range(k, k + m))
        else:
            wipe = rng.choice(k + m, int(rng.integers(1, m + 1)), replace=False).tolist()
            cur = full.copy()
            cur[wipe] = 0
            shards = rows_to_dev(cur)
            batch.add_reconstruct(codecs[(k, m)], shards, [i not in wipe for i in range(k + m)])
        cases.append((shards, full, wipe))
    batch.prepare()
    batch.set_graph(True)
    scratch = torch.zeros(1 << 20, dtype=torch.uint8, device="cuda")
    for replay in range(4):
        for shards, _, wipe in cases:
            for i in wipe:
                shards[i].zero_()
        scratch.add_(1)  # eager work between replays
        torch.cuda.synchronize()
        batch.run()
        torch.cuda.synchronize()
        for shards, full, _ in cases:
            assert (to_host(shards) == full).all(), replay
    batch.set_graph(False)
    for shards, _, wipe in cases:
        for i in wipe:
            shards[i].zero_()
    batch.run()
    torch.cuda.synchronize()
    for shards, full, _ in cases:
        assert (to_host(shards) == full).all()


def test_batch_fft_jobs_match_oracle(trs):
    """Shapes with k > 32 whose staged program or FFT decoder the batch runs
    on the byte-plane interpreter (one launch for all of them), mixed with
    matrix-route jobs, eager and as graph replays."""
    import torch

    rng = np.random.default_rng(33)
    codecs = {s: trs.ReedSolomon(*s) for s in [(64, 64), (100, 156), (128, 128), (48, 16)]}
    batch = trs.Batch()
    cases = []
    plan = [((64, 64), "enc", 4096), ((64, 64), "enc", 65536 + 5), ((100, 156), "enc", 1000),
            ((128, 128), "rec", 3000), ((128, 128), "rec", 777), ((48, 16), "enc", 5000),
            ((64, 64), "rec", 4096), ((128, 128), "rec-data", 2048)]
    for n, ((k, m), kind, length) in enumerate(plan):
        data = seeded(k, length, 900 + n)
        full = np.concatenate([data, O.rs8_encode(O.build_matrix(8, k, k + m), k, data)])
        if kind == "enc":
            shards = rows_to_dev(np.concatenate([data, np.zeros((m, length), np.uint8)]))
            batch.add_encode(codecs[(k, m)], shards)
            wipe, check = list(range(k, k + m)), list(range(k + m))
        else:
            e = 120 if n % 2 else 10  # 120 of 256: past the batched FFT-decoder threshold
            wipe = rng.choice(k + m, e, replace=False).tolist()
            cur = full.copy()
            cur[wipe] = 0
            shards = rows_to_dev(cur)
            data_only = kind == "rec-data"
            batch.add_reconstruct(codecs[(k, m)], shards, [i not in wipe for i in range(k + m)],
                                  data_only=data_only)
            check = [i for i in range(k + m) if not (data_only and i >= k and i in wipe)]
        cases.append((shards, full, wipe, check))
    batch.prepare()
    for graph in (False, True, True):
        batch.set_graph(graph)
        for shards, _, wipe, _ in cases:
            for i in wipe:
                shards[i].zero_()
        batch.run()
        torch.cuda.synchronize()
        for shards, full, _, check in cases:
            got = to_host(shards)
            assert (got[check] == full[check]).all(), graph


def test_batch_reprepare_waits_for_previous_run(trs):
    """clear -> add -> prepare right after run on a non-blocking side stream:
    prepare rewrites the device work list that the earlier run's kernels read,
    so it must wait for them; both batches' outputs stay exact."""
    import torch

    k, m, L = 10, 4, 1 << 20
    rs = trs.ReedSolomon(k, m)
    rs.set_apply_route("table")  # the runtime-coefficient list (k_batch8_planes)
    G = O.build_matrix(8, k, k + m)
    side = torch.cuda.Stream()
    sets = []
    for t in range(2):
        stripes = []
        for s in range(12):
            data = seeded(k, L, 7000 + 100 * t + s)
            stripes.append((rows_to_dev(np.concatenate([data, np.zeros((m, L), np.uint8)])),
                            data))
        sets.append(stripes)
    torch.cuda.synchronize()
    b = trs.Batch()
    for t, stripes in enumerate(sets):
        b.clear()
        for shards, _ in stripes:
            # reconstruct of every parity shard: the runtime-coefficient tile
            b.add_reconstruct(rs, shards, [True] * k + [False] * m)
        b.prepare()
        b.run(side.cuda_stream)
    side.synchronize()
    for stripes in sets:
        for shards, data in stripes:
            assert (to_host(shards[k:]) == O.rs8_encode(G, k, data)).all()
