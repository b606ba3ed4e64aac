"""The continuous-batching server: every request's result is byte-identical
to the same request run alone, while requests of mixed fields, shapes and
lengths overlap on the device and a tight HBM budget forces admission waits;
errors surface at submit exactly like the single calls."""
import threading

import numpy as np
import pytest

import oracle as O
from util import rows_to_dev, seeded, to_host

pytestmark = pytest.mark.gpu

SHAPES8 = [(10, 4), (7, 13), (10, 10), (32, 32), (18, 6)]
SHAPES16 = [(10, 4), (7, 13)]


def _requests(trs, n, seed):
    rng = np.random.default_rng(seed)
    codecs8 = {s: trs.ReedSolomon(*s) for s in SHAPES8}
    codecs16 = {s: trs.ReedSolomon16(*s) for s in SHAPES16}
    reqs = []
    for i in range(n):
        gf16 = i % 10 == 9
        k, m = (SHAPES16 if gf16 else SHAPES8)[int(rng.integers(0, 2 if gf16 else 5))]
        length = int(rng.choice([64, 4096, 12290, 65536, 200002]))
        data = seeded(k, length, 500 + i)
        rs = (codecs16 if gf16 else codecs8)[(k, m)]
        G = O.build_matrix(16 if gf16 else 8, k, k + m)
        parity = (O.rs16_apply(G[k:], list(data), length) if gf16 else O.rs8_encode(G, k, data))
        full = np.concatenate([data, parity])
        if rng.random() < 0.7:
            start = np.concatenate([data, np.zeros((m, length), np.uint8)])
            reqs.append(("enc", rs, start, None, full))
        else:
            e = int(rng.integers(1, m + 1))
            erased = set(rng.choice(k + m, e, replace=False).tolist())
            start = full.copy()
            for j in erased:
                start[j] = 0
            reqs.append(("rec", rs, start, [j not in erased for j in range(k + m)], full))
    return reqs


def test_server_mix_equals_alone(trs):
    import torch

    reqs = _requests(trs, 80, 3)
    # alone: each request through the plain per-call path
    alone = []
    for kind, rs, start, present, full in reqs:
        sh = rows_to_dev(start)
        if kind == "enc":
            rs.encode(sh)
        else:
            rs.reconstruct([(t, p) for t, p in zip(sh, present)])
        alone.append(to_host(sh))
        assert (alone[-1] == full).all()
    # in the mix: two submitter threads, 3 slots, a budget below the total size
    total = sum(r[2].nbytes for r in reqs)
    srv = trs.Server(slots=3, hbm_budget_bytes=total // 4, max_batch=16, window_us=20)
    bufs = [rows_to_dev(r[2]) for r in reqs]
    torch.cuda.synchronize()
    tickets = [None] * len(reqs)

    def submit(idx):
        for i in idx:
            kind, rs, _, present, _ = reqs[i]
            if kind == "enc":
                tickets[i] = srv.submit_encode(rs, bufs[i])
            else:
                tickets[i] = srv.submit_reconstruct(rs, bufs[i], present)

    th = [threading.Thread(target=submit, args=(range(j, len(reqs), 2),)) for j in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for t in tickets:
        srv.wait(t)
    for i, t in enumerate(tickets):
        lat, nb = srv.take(t)
        assert lat >= 0
        assert (to_host(bufs[i]) == alone[i]).all(), i
    srv.close()


def test_server_errors_at_submit(trs):
    rs = trs.ReedSolomon(4, 2)
    srv = trs.Server(slots=1)
    sh = rows_to_dev(np.zeros((5, 64), np.uint8))
    with pytest.raises(trs.Error) as e:
        srv.submit_encode(rs, sh)
    assert e.value.name == "TooFewShards"
    sh = rows_to_dev(np.zeros((6, 64), np.uint8))
    with pytest.raises(trs.Error) as e:
        srv.submit_reconstruct(rs, sh, [True, True, True, False, False, False])
    assert e.value.name == "TooFewShardsPresent"
    t = srv.submit_reconstruct(rs, sh, [True] * 6)  # all present: completes at once
    assert srv.poll(t)
    srv.close()
