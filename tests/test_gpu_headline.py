"""The headline batched workload at full size: 4096 stripes of the six
generated shapes, shard lengths log-uniform in [4 KiB, 1 MiB] (rounded to
64 B), erasures drawn from 64 fixed patterns per shape -- the same seeded
workload bench.py times.  Batched encode writes every parity shard; every
stripe then verifies through the per-call verify path (a different kernel
route), sampled stripes match the oracle, and the batched reconstruct (eager
and as graph replays) rebuilds every erased shard byte for byte."""
import os
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_full_batched_workload_is_byte_exact(trs):
    import torch

    sys.path.insert(0, ROOT)
    import bench

    rng = np.random.default_rng(bench.SEED)
    meta = bench.build_workload(4096, rng)
    assert len(meta) == 4096 and max(L for *_, L, _ in meta) <= 1 << 20
    codecs = {s: trs.ReedSolomon(*s) for s in bench.SHAPES}
    sizes = [(k + m) * L for (k, m, L, _) in meta]
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    buf = torch.zeros(int(offsets[-1]), dtype=torch.uint8, device="cuda")
    views = []
    enc, rec = trs.Batch(), trs.Batch()
    for s, (k, m, L, erased) in enumerate(meta):
        o = int(offsets[s])
        trs.fill_random(buf[o:o + k * L], bench.SEED + s)
        shards = [buf[o + i * L:o + (i + 1) * L] for i in range(k + m)]
        views.append(shards)
        enc.add_encode(codecs[(k, m)], shards)
        rec.add_reconstruct(codecs[(k, m)], shards, [i not in erased for i in range(k + m)])
    enc.run()
    torch.cuda.synchronize()
    pristine = buf.clone()

    # every stripe's parity, checked by the per-call verify route
    for s, (k, m, L, _) in enumerate(meta):
        assert codecs[(k, m)].verify(views[s]), s
    # sampled stripes against the oracle, one per 256 and the longest of each shape
    gens = {sh: O.build_matrix(8, sh[0], sh[0] + sh[1]) for sh in bench.SHAPES}
    longest = {}
    for s, (k, m, L, _) in enumerate(meta):
        if (k, m) not in longest or L > meta[longest[(k, m)]][2]:
            longest[(k, m)] = s
    for s in sorted(set(range(0, 4096, 256)) | set(longest.values())):
        k, m, L, _ = meta[s]
        o = int(offsets[s])
        data = O.fill(bench.SEED + s, k * L).reshape(k, L)
        got = pristine[o + k * L:o + (k + m) * L].cpu().numpy().reshape(m, L)
        assert (got == O.rs8_encode(gens[(k, m)], k, data)).all(), s

    def erase():
        for s, (k, m, L, erased) in enumerate(meta):
            o = int(offsets[s])
            for i in erased:
                buf[o + i * L:o + (i + 1) * L].zero_()

    rec.prepare()
    erased_bytes = sum(len(e) * L for (_, _, L, e) in meta)
    assert erased_bytes > 0 and rec.bytes() == sum((k + len(e)) * L for (k, _, L, e) in meta)
    for graph in (False, True, True):
        rec.set_graph(graph)
        erase()
        torch.cuda.synchronize()
        assert not torch.equal(buf, pristine)
        rec.run()
        torch.cuda.synchronize()
        assert torch.equal(buf, pristine), graph
