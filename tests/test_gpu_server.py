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


def test_server_admission_respects_free_hbm(trs):
    """Admission against free device memory: with a reserve above the device's
    free HBM a slot's table buffer may not grow, so a burst that queues behind
    a long request is split down to requests that run alone (as their plain
    per-call launch, which needs no tables); with the reserve lifted the same
    burst batches again.  Every request's parity equals the plain call's."""
    import ctypes

    import torch

    lib = trs._lib.lib
    k, m, L = 32, 32, 8 << 20
    rs = trs.ReedSolomon(k, m)
    buf = torch.zeros(((k + m), L), dtype=torch.uint8, device="cuda")
    trs.fill_random(buf[:k].reshape(-1), 0x5E7)
    shards = [buf[i] for i in range(k + m)]
    ref = buf.clone()
    rs.encode([ref[i] for i in range(k + m)])
    torch.cuda.synchronize()
    assert rs.verify([ref[i] for i in range(k + m)])
    arr = (ctypes.c_void_p * (k + m))(*[s.data_ptr() for s in shards])
    lens = (ctypes.c_size_t * (k + m))(*([L] * (k + m)))

    def burst(n):
        """n submissions of the stripe straight through the C ABI (a few
        microseconds each, so they queue behind the first, 512 MiB request)."""
        tickets = []
        for _ in range(n):
            tk = ctypes.c_uint64()
            assert lib.trs_server_submit_encode8(srv._h, rs._h, arr, lens, k + m,
                                                 ctypes.byref(tk)) == 0
            tickets.append(tk.value)
        first_running = not srv.poll(tickets[0])
        srv.drain()
        for tk in tickets:
            st = ctypes.c_int(0)
            assert lib.trs_server_wait(srv._h, tk, ctypes.byref(st)) == 0 and st.value == 0
            lat, ph = ctypes.c_double(0), (ctypes.c_double * 3)()
            assert lib.trs_server_take_detail(srv._h, tk, ctypes.byref(lat), ph, None,
                                              None) == 0
            assert min(ph) >= 0 and abs(sum(ph) - lat.value) < 1.0
        return first_running

    srv = trs.Server(slots=1, hbm_budget_bytes=0, max_batch=256, window_us=0)
    free, _ = torch.cuda.mem_get_info()
    srv.set_hbm_reserve(free + (1 << 40))
    buf[k:].zero_()
    torch.cuda.synchronize()
    queued = burst(16)
    s = srv.stats()
    assert s["done"] == 16 and s["oom"] == 0
    if queued:  # the other 15 were queued together: a 15-request batch was split
        assert s["splits"] >= 1 and s["batches"] == 16, s
    torch.cuda.synchronize()
    assert torch.equal(buf, ref)

    srv.set_hbm_reserve(0)  # room again: the slot's table buffer grows
    buf[k:].zero_()
    torch.cuda.synchronize()
    splits = s["splits"]
    queued = burst(16)
    s = srv.stats()
    assert s["done"] == 32 and s["oom"] == 0 and s["splits"] == splits
    if queued:
        assert s["batches"] < 32, s
    torch.cuda.synchronize()
    assert torch.equal(buf, ref)
    srv.close()


def test_server_spin_settings_keep_results(trs):
    """Blocking at once (spin 0) and polling give the same results; the phase
    breakdown adds up to the latency."""
    import torch

    reqs = _requests(trs, 24, 11)
    srv = trs.Server(slots=2, max_batch=8, window_us=0)
    for spin in (0, 500):
        srv.set_spin(spin)
        bufs = [rows_to_dev(r[2]) for r in reqs]
        torch.cuda.synchronize()
        tickets = []
        for (kind, rs, _, present, _), b in zip(reqs, bufs):
            tickets.append(srv.submit_encode(rs, b) if kind == "enc"
                           else srv.submit_reconstruct(rs, b, present))
        srv.drain()
        for (_, _, _, _, full), b, t in zip(reqs, bufs, tickets):
            lat, ph, _ = srv.take_detail(t)
            assert min(ph) >= 0 and abs(sum(ph) - lat) < 1.0
            assert (to_host(b) == full).all()
    with pytest.raises(trs.Error):
        srv.set_spin(200_000)
    srv.close()
