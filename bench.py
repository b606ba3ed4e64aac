#!/usr/bin/env python3
"""Benchmark: batched GF(2^8) encode + reconstruct on one MI355X.

Workload (the batched mixed-shape configuration): 4096 stripes drawn
uniformly from the shapes (7,13) (10,10) (14,14) (16,16) (18,6) (32,32),
shard length log-uniform in [4 KiB, 1 MiB] rounded to 64 B, uniform-random
bytes from a seeded stream; every stripe erases 1..m shards taken from a pool
of 64 fixed random erasure patterns per shape.  All shards live in HBM.

One step = one batched encode launch over all 4096 stripes followed by one
batched reconstruct launch over all 4096 stripes (decode plans cached in HBM).
`value` = algorithmic bytes of both launches / step time, where a stripe's
encode moves (k+m)*L bytes and its reconstruct (k+e)*L bytes (k survivor reads
plus e rebuilt writes).  The same work is also run as plain per-call GPU
encode / reconstruct (one library call per stripe) for comparison, with p50 /
p99 per-request latency.  Results are byte-checked: every reconstructed shard
against the pristine stripe, sampled parities against the CPU oracle.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
SHAPES = [(7, 13), (10, 10), (14, 14), (16, 16), (18, 6), (32, 32)]
SEED = 0x7A9E_5EED_0000_0002
HBM_PEAK_GBS = 8000.0


def build_workload(stripes: int, rng: np.random.Generator):
    """Shapes, lengths and erasure patterns of every stripe (host metadata)."""
    pools = {}
    for (k, m) in SHAPES:
        pool = []
        for _ in range(64):
            e = int(rng.integers(1, m + 1))
            pool.append(sorted(rng.choice(k + m, e, replace=False).tolist()))
        pools[(k, m)] = pool
    lo, hi = math.log(4096), math.log(1 << 20)
    meta = []
    for _ in range(stripes):
        k, m = SHAPES[int(rng.integers(0, len(SHAPES)))]
        L = int(round(math.exp(rng.uniform(lo, hi)) / 64.0)) * 64
        L = max(4096, min(1 << 20, L))
        erased = pools[(k, m)][int(rng.integers(0, 64))]
        meta.append((k, m, L, erased))
    return meta


SERVE_SHAPES = [(10, 4), (7, 13), (10, 10), (14, 14), (16, 16), (18, 6), (32, 32)]


def served_leg(trs, torch, requests: int, rates, seed: int, pool: int = 512):
    """Served mixed workload: Poisson arrivals, 70 % encode / 30 % reconstruct,
    shapes of the C1/C2 configurations, 90 % GF(2^8) / 10 % GF(2^16), shard
    length log-uniform [4 KiB, 4 MiB], erasures uniform in 1..m, device-resident
    shards.  `requests` requests per rate are drawn from a pool of `pool`
    encoded stripes (a request on a stripe rewrites bytes it already holds, so
    overlapping requests on one stripe are harmless); continuous batching
    (4 slots, 64 GiB budget) against the same requests issued as plain
    per-request calls.  Latency is split into queued / host launch / device
    phases, and the lowest rate is repeated with the server's threads polling
    for 200 us before they block (the default blocks at once).  The plain
    calls' own throughput bounds the rate they could sustain at all."""
    rng = np.random.default_rng(seed)
    c8 = {s: trs.ReedSolomon(*s) for s in SERVE_SHAPES}
    c16 = {}
    stripes = []
    lo, hi = math.log(4096), math.log(4 << 20)
    for i in range(pool):
        gf16 = rng.random() < 0.10
        k, m = SERVE_SHAPES[int(rng.integers(0, len(SERVE_SHAPES)))]
        L = int(round(math.exp(rng.uniform(lo, hi)) / 64.0)) * 64
        if gf16 and (k, m) not in c16:
            c16[(k, m)] = trs.ReedSolomon16(k, m)
        rs = c16[(k, m)] if gf16 else c8[(k, m)]
        buf = torch.empty((k + m) * L, dtype=torch.uint8, device="cuda")
        trs.fill_random(buf[:k * L], seed + i)
        shards = trs.pack_shards([buf[j * L:(j + 1) * L] for j in range(k + m)])
        rs.encode(shards)
        stripes.append((rs, shards, k, m, L))
    reqs = []
    for _ in range(requests):
        rs, shards, k, m, L = stripes[int(rng.integers(0, pool))]
        if rng.random() < 0.7:
            reqs.append(("enc", rs, shards, None, (k + m) * L))
        else:
            e = int(rng.integers(1, m + 1))
            erased = set(rng.choice(k + m, e, replace=False).tolist())
            reqs.append(("rec", rs, shards, [j not in erased for j in range(k + m)], (k + e) * L))
    torch.cuda.synchronize()
    total_bytes = sum(r[4] for r in reqs)

    # plain per-request calls, one after another
    stream = torch.cuda.current_stream()
    lat = []
    t0 = time.perf_counter()
    for kind, rs, shards, present, _ in reqs:
        a = time.perf_counter()
        if kind == "enc":
            rs.encode(shards)
        else:
            rs.reconstruct([(t, p) for t, p in zip(shards, present)])
        stream.synchronize()
        lat.append((time.perf_counter() - a) * 1e6)
    plain_s = time.perf_counter() - t0
    plain = {"gbs": round(total_bytes / plain_s / 1e9, 2),
             "p50_us": round(float(np.percentile(lat, 50)), 1),
             "p99_us": round(float(np.percentile(lat, 99)), 1),
             "max_req_s": round(len(reqs) / plain_s, 1)}

    out = {"requests_per_rate": requests, "pool_stripes": pool, "bytes": total_bytes,
           "plain_per_request": plain, "rates": []}
    srv = trs.Server(slots=4, hbm_budget_bytes=64 << 30, max_batch=256, window_us=50)

    def run(rate, n, spin):
        srv.set_spin(spin)
        sub = reqs[:n]
        sched = np.cumsum(rng.exponential(1.0 / rate, n))
        tickets, submit_t = [], []
        start = time.perf_counter()
        for (kind, rs, shards, present, _), at in zip(sub, sched):
            while time.perf_counter() - start < at:
                pass
            submit_t.append(time.perf_counter())
            if kind == "enc":
                tickets.append(srv.submit_encode(rs, shards))
            else:
                tickets.append(srv.submit_reconstruct(rs, shards, present))
        srv.drain()
        lats, phases, ends = [], [], []
        for t, st in zip(tickets, submit_t):
            l_us, ph, _ = srv.take_detail(t)
            lats.append(l_us)
            phases.append(ph)
            ends.append(st + l_us * 1e-6)
        span = max(ends) - min(submit_t)
        ph = np.asarray(phases)
        return {
            "rate_req_s": rate,
            "requests": n,
            "spin_us": spin,
            "gbs": round(sum(r[4] for r in sub) / span / 1e9, 2),
            "p50_us": round(float(np.percentile(lats, 50)), 1),
            "p99_us": round(float(np.percentile(lats, 99)), 1),
            "p50_phases_us": {"queued": round(float(np.median(ph[:, 0])), 1),
                              "launch": round(float(np.median(ph[:, 1])), 1),
                              "device": round(float(np.median(ph[:, 2])), 1)},
            "offered_s": round(float(sched[-1]), 4),
            "submitted_req_s": round(n / max(submit_t[-1] - submit_t[0], 1e-9), 1),
            "span_s": round(span, 4),
        }

    for rate in rates:
        out["rates"].append(run(rate, len(reqs), 0))
    out["polling_lowest_rate"] = run(rates[0], min(len(reqs), 3000), 200)
    out["server_stats"] = srv.stats()
    srv.close()
    return out


C1_SEED = 0x7A9E_5EED_0000_0001
TOWER_SEED = 0x7A9E_5EED_0000_0003


def host_leg(trs, torch, O, L: int = 32 << 20, reps: int = 3):
    """Host-resident shards (SURVEY (f) 2): RS8 (10,4) with 32 MiB shards in
    host memory, encode and a 4-erasure reconstruct through the pipelined host
    path, with pinned (hipHostMalloc, through torch) and pageable (numpy)
    buffers.  Rates count the stripe's moved bytes per second including both
    PCIe directions; results are checked against the device path's parity."""
    k, m = 10, 4
    rs = trs.ReedSolomon(k, m)
    dev = torch.empty((k + m, L), dtype=torch.uint8, device="cuda")
    trs.fill_random(dev[:k].reshape(-1), SEED ^ 0x405)
    rs.encode([dev[i] for i in range(k + m)])
    want = dev.cpu().numpy()
    del dev
    torch.cuda.empty_cache()
    out = {"workload": f"RS8 ({k},{m}), {L >> 20} MiB shards in host memory, 1 MiB pipeline chunks"}
    for kind in ("pinned", "pageable"):
        if kind == "pinned":
            rows = [torch.from_numpy(want[i].copy()).pin_memory() for i in range(k + m)]
            host = [r.numpy() for r in rows]
        else:
            rows = host = [want[i].copy() for i in range(k + m)]
        for j in range(k, k + m):
            host[j][:] = 0
        rs.encode_sep(rows[:k], rows[k:])
        t0 = time.perf_counter()
        for _ in range(reps):
            rs.encode_sep(rows[:k], rows[k:])
        enc_s = (time.perf_counter() - t0) / reps
        exact = all((host[j] == want[j]).all() for j in range(k, k + m))
        erased = [0, 3, 11, 13]
        present = [i not in erased for i in range(k + m)]
        t0 = time.perf_counter()
        for _ in range(reps):
            for i in erased:
                host[i][:] = 0
            rs.reconstruct([(r, p) for r, p in zip(rows, present)])
        rec_s = (time.perf_counter() - t0) / reps
        exact = exact and all((host[i] == want[i]).all() for i in range(k + m))
        out[kind] = {"encode_gbs": round((k + m) * L / enc_s / 1e9, 2),
                     "reconstruct_gbs": round((k + len(erased)) * L / rec_s / 1e9, 2),
                     "byte_exact": bool(exact)}
        del rows, host
    return out


def costmodel_leg(trs, n_train: int = 240, n_test: int = 60, seed: int = 0xC057):
    """Learned kernel choice (the cost-model configuration, scaled down to run
    in the default bench): shapes sampled as in the full study (field, k, m,
    log-uniform lengths 1 KiB - 4 MiB), every kernel variant of each timed on
    the GPU, the FP32 MLP trained on the GPU on the first n_train shapes; on the
    n_test held-out shapes, the regret against the exhaustive optimum of the
    model's measured top 3, of its top-1 pick, of the codec's own rules and of
    the reference's routing (both without a model)."""
    from reed_solomon_amd import autotune as A

    t0 = time.perf_counter()
    shapes = A.sample_shapes(n_train + n_test, seed)
    timings = [A.exhaustive(s) for s in shapes]
    t1 = time.perf_counter()
    X, T = A.dataset(shapes[:n_train], timings[:n_train])
    cm = A.CostModel(seed=1)
    loss = cm.fit(X, T, steps=1500)
    t2 = time.perf_counter()
    held, meas = shapes[n_train:], timings[n_train:]

    def stats(choose):
        r = np.array([A.regret(t, choose(s, t)) for s, t in zip(held, meas)])
        return {"mean": round(float(r.mean()), 4), "p90": round(float(np.percentile(r, 90)), 4),
                "within_10pct": round(float((r <= 0.10).mean()), 4)}

    return {
        "workload": f"{n_train + n_test} sampled shapes (GF(2^8)/GF(2^16), k, m, 1 KiB - 4 MiB), "
                    f"{n_test} held out; candidates = kernel variants",
        "train_samples": int(len(T)), "train_loss": round(float(loss), 5),
        "measure_s": round(t1 - t0, 1), "train_s": round(t2 - t1, 1),
        "candidates_per_shape": round(float(np.mean([len(t) for t in meas])), 2),
        "regret_model_top3": stats(lambda s, t: A.Tuner(cm, top_n=3).choose(s, t)),
        "regret_model_top1": stats(lambda s, t: A.Tuner(cm, top_n=1).choose(s, t)),
        "regret_codec_rules": stats(lambda s, t: A.rule(s)),
        "regret_reference_routing": stats(lambda s, t: A.heuristic(s)),
        "seed": seed,
    }


def wide_leg(trs, torch, reps: int = 20):
    """GF(2^8) shapes with k > 32 (SURVEY (f) 4 and (f) 1 over GF(2^8)): per
    call, the staged FFT program (src/fft.rs:393-592) against the matrix route
    for encode, and the FFT erasure decoder against the matrix plans for a
    reconstruct of m erasures; batched, a work list of such stripes on the
    interpreter list.  Each pair of routes must produce the same bytes."""

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    out, exact = {}, True
    for (k, m), L in (((64, 64), 64 << 10), ((200, 55), 64 << 10), ((128, 128), 1 << 20)):
        rs = trs.ReedSolomon(k, m)
        buf = torch.empty((k + m, L), dtype=torch.uint8, device="cuda")
        trs.fill_random(buf[:k].reshape(-1), SEED ^ (k << 8) ^ m)
        rows = [buf[i] for i in range(k + m)]
        enc = {}
        for route in ("table", "staged-fft"):
            ms = timed(lambda: rs.encode_sep_route(route, rows[:k], rows[k:]))
            enc[route] = (ms, buf[k:].clone())
        exact = exact and bool(torch.equal(enc["table"][1], enc["staged-fft"][1]))
        full = buf.clone()
        erased = list(range(0, k + m, (k + m) // m))[:m]
        present = [i not in erased for i in range(k + m)]
        rec = {}
        lib = trs._lib.lib
        p, l = trs._ptrs(rows)
        flags = trs._flags(present)
        for route in ("matrix", "fft"):
            rs.set_decode_route(route)
            dec = rs.prepare_decode(present)

            def run():
                st = lib.trs_dec8_reconstruct(dec._h, p, l, flags, k + m, 0, trs._stream())
                assert st == 0, st

            rec[route] = timed(run)
            for i in erased:
                rows[i].zero_()
            run()
            torch.cuda.synchronize()
            exact = exact and bool(torch.equal(buf, full))
        rs.set_decode_route("auto")
        mb = (k + m) * L
        out[f"{k}_{m}_{L >> 10}KiB"] = {
            "encode_gbs": {"matrix": round(mb / (enc["table"][0] * 1e6), 1),
                           "staged_fft": round(mb / (enc["staged-fft"][0] * 1e6), 1)},
            "encode_route_auto": rs.encode_route(L),
            "encode_route_reference": rs.reference_route(L),
            "reconstruct_gbs": {"matrix": round((k + m) * L / (rec["matrix"] * 1e6), 1),
                                "fft_decoder": round((k + m) * L / (rec["fft"] * 1e6), 1)},
        }
        del buf, rows, full
        torch.cuda.empty_cache()
    # batched: 256 stripes of (64,64) x 64 KiB through the interpreter work list
    k, m, L, S = 64, 64, 64 << 10, 256
    rs = trs.ReedSolomon(k, m)
    buf = torch.empty((S, k + m, L), dtype=torch.uint8, device="cuda")
    trs.fill_random(buf[:, :k].reshape(-1), SEED ^ 0x51DE)
    b = trs.Batch()
    for st in range(S):
        b.add_encode(rs, [buf[st, i] for i in range(k + m)])
    b.prepare()
    b.set_graph(True)
    ms = timed(b.run)
    one = [buf[7, i].clone() for i in range(k + m)]
    rs.encode_sep_route("table", one[:k], one[k:])
    exact = exact and all(bool(torch.equal(one[i], buf[7, i])) for i in range(k, k + m))
    out["batched_64_64_x256_64KiB_encode_gbs"] = round(S * (k + m) * L / (ms * 1e6), 1)
    out["byte_exact"] = exact
    del buf, b
    torch.cuda.empty_cache()
    return out


def c1_leg(trs, torch, O, stripes: int = 256, L: int = 65536):
    """RS8 (10,4), `stripes` stripes of 14 x L bytes resident in HBM, seeded
    bytes: batched encode, then per stripe erase 4 of 14 shards chosen
    uniformly (any data/parity mix) and batched reconstruct; byte-compared with
    the pristine stripes and, on sampled stripes, the oracle's parity.  Also
    times batched encode of the same shape at larger shards (the >= 256 KiB
    regime) as a fraction of HBM peak."""
    k, m = 10, 4
    n = k + m
    rs = trs.ReedSolomon(k, m)
    buf = torch.empty((stripes, n, L), dtype=torch.uint8, device="cuda")
    for s in range(stripes):
        trs.fill_random(buf[s, :k].reshape(-1), C1_SEED + s)
    rng = np.random.default_rng(C1_SEED)
    erased = [sorted(rng.choice(n, 4, replace=False).tolist()) for _ in range(stripes)]
    enc, rec = trs.Batch(), trs.Batch()
    for s in range(stripes):
        shards = [buf[s, i] for i in range(n)]
        enc.add_encode(rs, shards)
        rec.add_reconstruct(rs, shards, [i not in erased[s] for i in range(n)])
    enc.prepare()
    rec.prepare()
    enc.set_graph(True)
    rec.set_graph(True)
    enc.run()
    torch.cuda.synchronize()
    pristine = buf.clone()
    G = O.build_matrix(8, k, n)
    oracle_ok = all(bool((buf[s, k:].cpu().numpy() ==
                          O.rs8_encode(G, k, buf[s, :k].cpu().numpy())).all())
                    for s in (0, stripes // 2, stripes - 1))
    mask = torch.zeros((stripes, n), dtype=torch.bool, device="cuda")
    for s in range(stripes):
        mask[s, erased[s]] = True

    def timed(batch, before=None, reps=10):
        ev = []
        for _ in range(reps):
            if before is not None:
                before()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            batch.run()
            e1.record()
            ev.append((e0, e1))
        torch.cuda.synchronize()
        return float(np.median([a.elapsed_time(b) for a, b in ev]))

    enc_ms = timed(enc)
    rec_ms = timed(rec, lambda: buf.masked_fill_(mask[:, :, None], 0))
    exact = bool(torch.equal(buf, pristine))
    enc_bytes = stripes * n * L
    rec_bytes = stripes * (k + 4) * L
    del buf, pristine, enc, rec
    torch.cuda.empty_cache()
    large = []
    for Lb, S in ((256 << 10, 256), (1 << 20, 128), (4 << 20, 32)):
        t = torch.empty((S, n, Lb), dtype=torch.uint8, device="cuda")
        trs.fill_random(t.reshape(-1), C1_SEED ^ Lb)
        b = trs.Batch()
        for s in range(S):
            b.add_encode(rs, [t[s, i] for i in range(n)])
        b.prepare()
        b.set_graph(True)
        b.run()
        ms = timed(b)
        gbs = S * n * Lb / ms / 1e6
        large.append({"shard_bytes": Lb, "stripes": S, "ms": round(ms, 3), "gbs": round(gbs, 1),
                      "hbm_frac": round(gbs / HBM_PEAK_GBS, 3)})
        del t, b
        torch.cuda.empty_cache()
    return {
        "workload": f"RS8 (10,4), {stripes} stripes of 14 x {L // 1024} KiB; batched encode, then "
                    "4 uniformly erased shards per stripe and batched reconstruct",
        "encode_bytes": enc_bytes, "encode_ms": round(enc_ms, 3),
        "encode_gbs": round(enc_bytes / enc_ms / 1e6, 1),
        "reconstruct_bytes": rec_bytes, "reconstruct_ms": round(rec_ms, 3),
        "reconstruct_gbs": round(rec_bytes / rec_ms / 1e6, 1),
        "large_shard_encode": large,
        "byte_exact": {"reconstruct": exact, "oracle_sampled_parity": oracle_ok},
    }
def host_threads() -> int:
    """Host cores this process may use: its CPU affinity, capped by
    OMP_NUM_THREADS when that is set (the GPU boxes give one GPU's process a
    16-core share and set it to 16)."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def _run_parallel(tasks, threads: int) -> float:
    """Wall seconds to run `tasks` ((work_bytes, [callables]) pairs) on
    `threads` threads, largest first so that the threads finish together.
    ctypes releases the GIL inside the C calls, so the threads run them in
    parallel."""
    from concurrent.futures import ThreadPoolExecutor

    order = sorted(tasks, key=lambda t: -t[0])

    def one(task):
        for call in task[1]:
            call()

    a = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        list(ex.map(one, order))
    return time.perf_counter() - a


def cpu_baseline(O, meta, gens, seconds: float, sample_bytes: int = 2 << 30):
    """CPU encode + reconstruct of the first stripes of the workload.

    With GFNI and AVX-512 (the GPU box's EPYC has both) it times the
    reference's x86 algorithms restated in C (oracle/simd_ref.c and the
    straight-line encoders written from the generated op lists): encode through
    the generated LCH program of each shape (src/gf/fft_x86.rs:22-165),
    reconstruct as the two fused GFNI applies of a cached decode plan
    (src/gf/x86_fused.rs, src/reedsolomon.rs:829-891).  The sample is the first
    stripes up to `sample_bytes` of stripe memory; every stripe is checked
    (the first ones against the scalar oracle, all of them by the round trip)
    before timing.  Timed calls are single C calls with their pointer arrays
    built beforehand, repeated over the sample for about `seconds` on one
    thread, then on every host core.  Without GFNI / AVX-512 the scalar oracle
    runs the same stripes once."""
    simd = O.simd_available()
    threads = host_threads()
    tasks, plans = [], {}
    pass_bytes, held, nst = 0, 0, 0
    for s, (k, m, L, erased) in enumerate(meta):
        if held + 3 * (k + m) * L > sample_bytes and nst:
            break
        G = gens[(k, m)]
        data = O.fill(SEED + s, k * L).reshape(k, L)
        present = [i not in erased for i in range(k + m)]
        work = (k + m) * L + (k + len(erased)) * L
        if simd:
            parity = np.zeros((m, L), np.uint8)
            enc = O.simd_encode_call(G, k, m, data, parity)
            enc()
            if nst < 4 and not (parity == O.rs8_encode(G, k, data)).all():
                raise RuntimeError(f"GFNI encode differs from the oracle on stripe {s}")
            full = np.concatenate([data, parity])
            cur = full.copy()
            for i in erased:
                cur[i] = 0
            key = (k, m, tuple(erased))
            if key not in plans:
                plans[key] = O.rs8_decode_plan(G, k, m, present)
            rec = O.simd_reconstruct_calls(G, k, m, cur, plans[key])
            for call in rec:
                call()
            if not (cur == full).all():
                raise RuntimeError(f"GFNI reconstruct differs from the original on stripe {s}")
            tasks.append((work, [enc] + rec))
        else:
            cur = np.concatenate([data, O.rs8_encode(G, k, data)])
            for i in erased:
                cur[i] = 0
            tasks.append((work, [lambda G=G, k=k, d=data: O.rs8_encode(G, k, d),
                                 lambda G=G, k=k, m=m, c=cur, p=present:
                                 O.rs8_reconstruct(G, k, m, c, p)]))
        pass_bytes += work
        held += 3 * (k + m) * L
        nst += 1
    # one thread: whole passes over the sample until `seconds` have passed
    passes, spent = 0, 0.0
    while passes == 0 or (simd and spent < seconds):
        a = time.perf_counter()
        for _, calls in tasks:
            for call in calls:
                call()
        spent += time.perf_counter() - a
        passes += 1
    what = ("reference x86 algorithms restated in C: generated LCH programs (encode) and fused "
            "GFNI applies with cached plans (reconstruct), AVX-512" if simd else
            "scalar C oracle (MUL-table loop); this CPU lacks GFNI / AVX-512")
    out = {"value": round(passes * pass_bytes / spent / 1e9, 4), "unit": "GB/s", "cores": 1,
           "kind": "port",
           "sample": f"first {nst} stripes of the workload ({pass_bytes / 1e9:.2f} GB of "
                     f"encode + reconstruct bytes per pass), {what}; {passes} passes, "
                     f"{spent:.1f} s"}
    passes_all, wall = 0, 0.0
    while passes_all == 0 or (simd and wall < seconds / 4):
        wall += _run_parallel(tasks, threads)
        passes_all += 1
    out["all_cores"] = {"value": round(passes_all * pass_bytes / wall / 1e9, 4), "unit": "GB/s",
                        "cores": threads,
                        "sample": f"the same stripes, one stripe per task on {threads} threads, "
                                  f"{passes_all} passes, {wall:.2f} s wall"}
    return out


TRAFFIC_FILE = "profiles/hbm_traffic.json"


def traffic_for(kernel: str, algorithmic_bytes: int):
    """HBM bytes per launch of `kernel` on this workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs (scaled by a copy kernel of known
    size); None when that file does not cover the workload."""
    try:
        with open(os.path.join(ROOT, TRAFFIC_FILE)) as f:
            rec = json.load(f)["kernels"][kernel]
    except (OSError, KeyError, ValueError):
        return None
    if rec.get("algorithmic_bytes") != algorithmic_bytes:
        return None
    return rec["hbm_bytes_per_launch"]


def tower_leg(trs, torch, O, stripes: int = 64):
    """RS16 (1024,1024), 4 KiB split-plane shards (2048 symbols), `stripes`
    stripes resident in HBM: encode through the tower-FFT route and through the
    matrix route (one multi-stripe call each), then per stripe erase 1024
    uniformly chosen shards and reconstruct all stripes in one call with a
    decode plan per pattern (first call builds them, timed passes reuse them),
    through the additive-FFT erasure decoder (one launch for all stripes) and
    through the matrix route (one apply per stripe)."""
    k = m = 1024
    n, L = k + m, 4096
    rs = trs.ReedSolomon16(k, m)
    buf = torch.empty((stripes, n, L), dtype=torch.uint8, device="cuda")
    for s in range(stripes):
        trs.fill_random(buf[s, :k].reshape(-1), TOWER_SEED + s)
    enc_bytes = stripes * n * L

    def timed(route, reps):
        rs.encode_stripes(buf, route)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            rs.encode_stripes(buf, route)
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    fft_ms = timed("fft", 20)
    fft_parity = buf[:, k:].clone()
    matrix_ms = timed("matrix", 2)
    routes_equal = bool(torch.equal(fft_parity, buf[:, k:]))
    del fft_parity
    # oracle on sampled symbol columns of the first and last stripe
    G = O.build_matrix(16, k, n)
    cols = np.r_[0:24, 1000:1008, 2040:2048]
    oracle_ok = True
    for s in (0, stripes - 1):
        host = buf[s].cpu().numpy()
        sub = np.ascontiguousarray(np.concatenate([host[:k][:, cols], host[:k][:, L // 2 + cols]], 1))
        want = O.rs16_symbolwise(G[k:], list(sub), sub.shape[1])
        got = np.concatenate([host[k:][:, cols], host[k:][:, L // 2 + cols]], 1)
        oracle_ok &= bool((got == want).all())
    pristine = buf.clone()

    rng = np.random.default_rng(TOWER_SEED)
    present = np.ones((stripes, n), bool)
    for s in range(stripes):
        present[s, rng.choice(n, 1024, replace=False)] = False
    erased_mask = torch.from_numpy(~present).to("cuda")
    rec_bytes = stripes * (k + 1024) * L

    def erase_all():
        buf[erased_mask] = 0

    def reconstruct_all():
        rs.reconstruct_stripes(buf, present)

    rec = {}
    for route in ("fft", "matrix"):
        rs.set_decode_route(route)
        erase_all()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        reconstruct_all()  # builds (first route) and caches one plan per pattern
        torch.cuda.synchronize()
        first_s = time.perf_counter() - t0
        first_exact = bool(torch.equal(buf, pristine))
        ev = []
        for _ in range(5 if route == "fft" else 2):
            erase_all()
            # the masked fill returns before its kernels finish: wait, so
            # that the interval holds the call (host work + kernels) alone
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            reconstruct_all()
            e1.record()
            ev.append((e0, e1))
        torch.cuda.synchronize()
        ms = float(np.median([a.elapsed_time(b) for a, b in ev]))
        rec[route] = {"ms": round(ms, 3), "gbs": round(rec_bytes / ms / 1e6, 2),
                      "first_call_s": round(first_s, 3), "first_exact": first_exact,
                      "cached_exact": bool(torch.equal(buf, pristine))}
    del buf, pristine
    torch.cuda.empty_cache()
    return {
        "workload": f"RS16 (1024,1024), {stripes} stripes of 2048 x 4 KiB split-plane shards; "
                    "encode via tower FFT and via matrix; reconstruct of 1024 uniformly erased "
                    "shards per stripe (one pattern per stripe) via the FFT decoder and via "
                    "the matrix route",
        "encode_bytes": enc_bytes,
        "encode_fft": {"ms": round(fft_ms, 3), "gbs": round(enc_bytes / fft_ms / 1e6, 1)},
        "encode_matrix": {"ms": round(matrix_ms, 3), "gbs": round(enc_bytes / matrix_ms / 1e6, 2)},
        "fft_vs_matrix": round(matrix_ms / fft_ms, 1),
        "reconstruct_bytes": rec_bytes,
        "reconstruct_fft": {key: v for key, v in rec["fft"].items() if "exact" not in key},
        "reconstruct_matrix": {key: v for key, v in rec["matrix"].items() if "exact" not in key},
        "reconstruct_fft_vs_matrix": round(rec["matrix"]["ms"] / rec["fft"]["ms"], 1),
        "byte_exact": {"fft_equals_matrix": routes_equal, "oracle_sampled_columns": oracle_ok,
                       **{f"reconstruct_{r}_{w}": rec[r][f"{w}_exact"]
                          for r in ("fft", "matrix") for w in ("first", "cached")}},
    }


def mfma_run(trs, torch, stripes: int = 1024) -> int:
    """The int8-MFMA kernels beside the bit-sliced planes kernel on the same
    work, for the MFMA counter runs (SQ_INSTS_VALU_MFMA_MOPS_I8,
    SQ_VALU_MFMA_BUSY_CYCLES): per call, reconstructs of 3 data shards of RS8
    (10,4) and of 8 of (32,32) at 256 KiB (k_gf8_mfma against k_gf8_planes),
    and tower (10,4) and (16,16) encodes at 1 MiB (k_gf8_mfma on the byte
    problem against k_gf16_apply_inline);
    batched, the mixed reconstruct work list of the first `stripes` stripes of
    the default workload (k_batch_mfma against k_batch8_planes).  Prints the
    HIP-event time and algorithmic bytes of each."""
    out = {}

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    L = 256 << 10
    for (k, m, e) in ((10, 4, 3), (32, 32, 8)):
        rs = trs.ReedSolomon(k, m)
        buf = torch.empty((k + m, L), dtype=torch.uint8, device="cuda")
        trs.fill_random(buf[:k].reshape(-1), SEED ^ k)
        shards = [buf[i] for i in range(k + m)]
        rs.encode(shards)
        present = [i >= e for i in range(k + m)]
        pairs = [(t, p) for t, p in zip(shards, present)]
        for route in ("mfma", "table"):  # "table": the planes tile at this length
            rs.set_apply_route(route)
            ms = timed(lambda: rs.reconstruct(pairs), 50)
            out[f"percall_{k}_{m}_{route}"] = {"ms": round(ms, 4), "bytes": (k + e) * L,
                                               "gbs": round((k + e) * L / ms / 1e6, 1)}
        del buf
    # the tower: split-plane (10,4) and (16,16) encodes at 1 MiB on int8 MFMA
    # (the 16m x 16k bit expansion) against the Karatsuba kernel
    Lt = 1 << 20
    for (k, m) in ((10, 4), (16, 16)):
        rs = trs.ReedSolomon16(k, m)
        buf = torch.empty((k + m, Lt), dtype=torch.uint8, device="cuda")
        trs.fill_random(buf[:k].reshape(-1), SEED ^ (k << 8))
        data, parity = [buf[i] for i in range(k)], [buf[k + j] for j in range(m)]
        for route in ("mfma", "table"):
            rs.set_apply_route(route)
            ms = timed(lambda: rs.encode_sep_matrix(data, parity), 50)
            out[f"tower_percall_{k}_{m}_{route}"] = {"ms": round(ms, 4), "bytes": (k + m) * Lt,
                                                     "gbs": round((k + m) * Lt / ms / 1e6, 1)}
        del buf
    meta = build_workload(stripes, np.random.default_rng(SEED))
    sizes = [(k + m) * Ls for (k, m, Ls, _) in meta]
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    buf = torch.empty(int(offsets[-1]), dtype=torch.uint8, device="cuda")
    for route in ("mfma", "auto"):
        codecs = {sh: trs.ReedSolomon(*sh) for sh in SHAPES}
        for c in codecs.values():
            c.set_apply_route(route)
        b = trs.Batch()
        for s, (k, m, Ls, erased) in enumerate(meta):
            o = int(offsets[s])
            if route == "mfma":
                trs.fill_random(buf[o:o + k * Ls], SEED + s)
            shards = [buf[o + i * Ls:o + (i + 1) * Ls] for i in range(k + m)]
            if route == "mfma":
                codecs[(k, m)].encode(shards)
            b.add_reconstruct(codecs[(k, m)], shards, [i not in erased for i in range(k + m)])
        b.prepare()
        ms = timed(b.run, 5)
        out[f"batched_{route}"] = {"ms": round(ms, 3), "bytes": b.bytes(),
                                   "gbs": round(b.bytes() / ms / 1e6, 1)}
    print(json.dumps(out), flush=True)
    return 0


def counter_run(trs, torch, stripes: int) -> int:
    """The launches whose HBM traffic `roofline.traffic` reports -- the
    batched encode (k_batch_bs_*) and reconstruct (k_batch8_planes) work lists of the
    default workload, the tower FFT encode (k_fft16) and FFT decode
    (k_fft16_jobs) of the tower workload, three launches each -- then a
    calibration kernel of known traffic (bitwise_not over 1 GiB: 1 GiB read +
    1 GiB written)."""
    lib = trs._lib.lib
    sptr = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    meta = build_workload(stripes, np.random.default_rng(SEED))
    codecs = {s: trs.ReedSolomon(*s) for s in SHAPES}
    sizes = [(k + m) * L for (k, m, L, _) in meta]
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    buf = torch.empty(int(offsets[-1]), dtype=torch.uint8, device="cuda")
    base = buf.data_ptr()
    be, br = trs.Batch(), trs.Batch()
    for s, (k, m, L, erased) in enumerate(meta):
        o, n = int(offsets[s]), k + m
        trs.fill_random(buf[o:o + k * L], SEED + s)
        ptrs = (ctypes.c_void_p * n)(*[base + o + i * L for i in range(n)])
        lens = (ctypes.c_size_t * n)(*([L] * n))
        present = (ctypes.c_uint8 * n)(*[0 if i in erased else 1 for i in range(n)])
        rs = codecs[(k, m)]
        assert lib.trs_rs8_encode(rs._h, ptrs, lens, n, sptr) == 0
        assert lib.trs_batch8_add_encode(be._h, rs._h, ptrs, lens, n) == 0
        assert lib.trs_batch8_add_reconstruct(br._h, rs._h, ptrs, lens, present, n, 0) == 0
    be.prepare()
    br.prepare()
    for _ in range(3):
        be.run()
        br.run()
    torch.cuda.synchronize()
    print(json.dumps({"k_batch_bs": be.bytes(), "k_batch8_planes": br.bytes()}), flush=True)
    del buf
    torch.cuda.empty_cache()
    r16 = trs.ReedSolomon16(1024, 1024)
    t = torch.zeros((64, 2048, 4096), dtype=torch.uint8, device="cuda")
    trs.fill_random(t.reshape(-1), TOWER_SEED)
    for _ in range(3):
        r16.encode_stripes(t, "fft")
    rng = np.random.default_rng(TOWER_SEED)
    present = np.ones((64, 2048), bool)
    for s in range(64):
        present[s, rng.choice(2048, 1024, replace=False)] = False
    r16.set_decode_route("fft")
    for _ in range(3):
        r16.reconstruct_stripes(t, present)
    torch.cuda.synchronize()
    del t
    torch.cuda.empty_cache()
    a = torch.full((1 << 30,), 7, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        torch.bitwise_not(a, out=b)
    torch.cuda.synchronize()
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--stripes", type=int, default=4096)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=10000)
    ap.add_argument("--no-tower", action="store_true")
    ap.add_argument("--no-costmodel", action="store_true")
    ap.add_argument("--mfma", action="store_true",
                    help="run the int8-MFMA kernels beside the planes kernel (for counter runs)")
    ap.add_argument("--no-c1", action="store_true")
    ap.add_argument("--no-wide", action="store_true",
                    help="skip the k > 32 leg (staged FFT and FFT decoder vs matrix)")
    ap.add_argument("--no-host", action="store_true",
                    help="skip the host-resident (PCIe-inclusive) leg")
    ap.add_argument("--counters", action="store_true",
                    help="run only the measured launches (3 each) and a 1 GiB calibration "
                         "kernel, for rocprofv3 --pmc passes")
    ap.add_argument("--tower-stripes", type=int, default=64)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "across GPUs the stripes would shard with no exchange, each GPU running this "
              "path on its share", file=sys.stderr)
        return 2

    import torch

    sys.path.insert(0, ROOT)
    import __graft_entry__

    trs = __graft_entry__.load()
    lib = trs._lib.lib
    torch.cuda.set_device(0)
    stream = torch.cuda.current_stream()
    sptr = ctypes.c_void_p(stream.cuda_stream)

    if args.mfma:
        return mfma_run(trs, torch)
    if args.counters:
        return counter_run(trs, torch, args.stripes)

    rng = np.random.default_rng(SEED)
    meta = build_workload(args.stripes, rng)
    codecs = {s: trs.ReedSolomon(*s) for s in SHAPES}
    sizes = [(k + m) * L for (k, m, L, _) in meta]
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    total_bytes = int(offsets[-1])
    buf = torch.empty(total_bytes, dtype=torch.uint8, device="cuda")
    base = buf.data_ptr()

    # seeded data, then per-call encode to build every stripe
    enc_args, rec_args = [], []
    for s, (k, m, L, erased) in enumerate(meta):
        o = int(offsets[s])
        trs.fill_random(buf[o:o + k * L], SEED + s)
        n = k + m
        ptrs = (ctypes.c_void_p * n)(*[base + o + i * L for i in range(n)])
        lens = (ctypes.c_size_t * n)(*([L] * n))
        present = (ctypes.c_uint8 * n)(*[0 if i in erased else 1 for i in range(n)])
        rs = codecs[(k, m)]
        enc_args.append((rs._h, ptrs, lens, n))
        rec_args.append((rs._h, ptrs, lens, present, n))
        assert lib.trs_rs8_encode(rs._h, ptrs, lens, n, sptr) == 0
    torch.cuda.synchronize()
    pristine = buf.clone()

    # CPU oracle check of sampled parities (the checker, not the measured path)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    gens = {s: O.build_matrix(8, s[0], s[0] + s[1]) for s in SHAPES}
    sample = list(range(0, args.stripes, max(1, args.stripes // 12)))
    host_ok = True
    for s in sample:
        k, m, L, _ = meta[s]
        data = O.fill(SEED + s, k * L).reshape(k, L)
        o = int(offsets[s])
        got = pristine[o + k * L:o + (k + m) * L].cpu().numpy().reshape(m, L)
        host_ok &= bool((got == O.rs8_encode(gens[(k, m)], k, data)).all())

    def erase():
        for s, (k, m, L, erased) in enumerate(meta):
            o = int(offsets[s])
            for i in erased:
                buf[o + i * L:o + (i + 1) * L].zero_()

    # batched work lists
    batch_enc, batch_rec = trs.Batch(), trs.Batch()
    for (h, ptrs, lens, n) in enc_args:
        assert lib.trs_batch8_add_encode(batch_enc._h, h, ptrs, lens, n) == 0
    for (h, ptrs, lens, present, n) in rec_args:
        assert lib.trs_batch8_add_reconstruct(batch_rec._h, h, ptrs, lens, present, n, 0) == 0
    batch_enc.prepare()
    batch_rec.prepare()
    enc_bytes, rec_bytes = batch_enc.bytes(), batch_rec.bytes()

    # correctness of the batched reconstruct from a freshly erased state
    erase()
    batch_rec.run()
    torch.cuda.synchronize()
    rec_exact = bool(torch.equal(buf, pristine))
    # correctness of the batched encode: wipe all parity and rebuild it
    for s, (k, m, L, _) in enumerate(meta):
        o = int(offsets[s])
        buf[o + k * L:o + (k + m) * L].zero_()
    batch_enc.run()
    torch.cuda.synchronize()
    enc_exact = bool(torch.equal(buf, pristine))

    # ---- batched steps: each work list replays as one captured hipGraph
    batch_enc.set_graph(True)
    batch_rec.set_graph(True)

    def step_batched():
        batch_enc.run()
        batch_rec.run()

    for _ in range(args.warmup):
        step_batched()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True),
           torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for e0, e1, e2 in ev:
        e0.record()
        batch_enc.run()
        e1.record()
        batch_rec.run()
        e2.record()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / args.steps
    enc_ms = float(np.mean([a.elapsed_time(b) for a, b, _ in ev]))
    rec_ms = float(np.mean([b.elapsed_time(c) for _, b, c in ev]))
    value = (enc_bytes + rec_bytes) / (ms_step * 1e-3) / 1e9
    # dominant kernel: the batched reconstruct launch (k_batch8_planes)
    achieved = rec_bytes / (rec_ms * 1e-3) / 1e9
    # The timed replays rewrite bytes that are already correct, so check that
    # the same graphs really write: outside the timed region, wipe all parity
    # and replay the encode graph, then erase and replay the reconstruct graph.
    torch.cuda.synchronize()
    for s, (k, m, L, _) in enumerate(meta):
        o = int(offsets[s])
        buf[o + k * L:o + (k + m) * L].zero_()
    batch_enc.run()
    torch.cuda.synchronize()
    exact_after = bool(torch.equal(buf, pristine))
    erase()
    batch_rec.run()
    torch.cuda.synchronize()
    exact_after = exact_after and bool(torch.equal(buf, pristine))

    # ---- plain per-call GPU path: one library call per stripe
    def percall(kind: str):
        if kind == "enc":
            for (h, ptrs, lens, n) in enc_args:
                lib.trs_rs8_encode(h, ptrs, lens, n, sptr)
        else:
            for (h, ptrs, lens, present, n) in rec_args:
                lib.trs_rs8_reconstruct(h, ptrs, lens, present, n, 0, sptr)

    percall("enc")
    percall("rec")
    torch.cuda.synchronize()
    reps = max(1, args.steps // 5)
    t0 = time.perf_counter()
    for _ in range(reps):
        percall("enc")
    torch.cuda.synchronize()
    t_enc = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        percall("rec")
    torch.cuda.synchronize()
    t_rec = (time.perf_counter() - t0) / reps
    percall_gbs = (enc_bytes + rec_bytes) / (t_enc + t_rec) / 1e9

    # per-request latency (submit -> complete), per-call reconstruct
    lat = []
    for (h, ptrs, lens, present, n) in rec_args[:1024]:
        a = time.perf_counter()
        lib.trs_rs8_reconstruct(h, ptrs, lens, present, n, 0, sptr)
        stream.synchronize()
        lat.append((time.perf_counter() - a) * 1e6)
    lat = np.array(lat)
    exact_percall = bool(torch.equal(buf, pristine))

    # ---- CPU baseline on a bounded sample of the workload, one thread and
    # every host core: the reference's GFNI / AVX-512 algorithms restated in C
    # where the CPU has them, the scalar oracle otherwise
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(O, meta, gens, args.cpu_seconds)

    del buf, pristine
    torch.cuda.empty_cache()
    tower = None
    if not args.no_tower:
        tower = tower_leg(trs, torch, O, args.tower_stripes)
    c1 = None
    if not args.no_c1:
        c1 = c1_leg(trs, torch, O)
    host = None
    if not args.no_host:
        host = host_leg(trs, torch, O)
    wide = None
    if not args.no_wide:
        wide = wide_leg(trs, torch)

    costmodel = None
    if not args.no_costmodel:
        costmodel = costmodel_leg(trs)

    served = None
    if not args.no_serve:
        served = served_leg(trs, torch, args.serve_requests, [1000, 5000, 20000, 50000],
                            SEED + 5)

    out = {
        "metric": "GB/s encode+reconstruct (byte-exact), batched vs per-call on 1 GPU; p99 latency",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (seeded splitmix64 bytes)",
        "config": {
            "workload": "batched GF(2^8) encode+reconstruct, 4096 stripes of shapes "
                        "(7,13),(10,10),(14,14),(16,16),(18,6),(32,32), shard length log-uniform "
                        "[4 KiB, 1 MiB] rounded to 64 B, erasures 1..m from 64 patterns per shape",
            "stripes": args.stripes,
            "bytes_per_step": enc_bytes + rec_bytes,
            "encode_bytes": enc_bytes,
            "reconstruct_bytes": rec_bytes,
            "resident_bytes": total_bytes,
            "seed": SEED,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic_for("k_batch8_planes", rec_bytes),
            "kernel": "k_batch8_planes",
            "algorithmic_bytes": rec_bytes,
            "kernel_ms": round(rec_ms, 3),
            "traffic_source": TRAFFIC_FILE,
            "other": {"k_batch_bs": {"role": "batched encode", "kernel_ms": round(enc_ms, 3),
                                     "achieved": round(enc_bytes / (enc_ms * 1e-3) / 1e9, 1),
                                     "traffic": traffic_for("k_batch_bs", enc_bytes)}},
        },
        "cpu_baseline": cpu,
        "batched": {
            "encode_gbs": round(enc_bytes / (enc_ms * 1e-3) / 1e9, 1),
            "reconstruct_gbs": round(rec_bytes / (rec_ms * 1e-3) / 1e9, 1),
        },
        "per_call": {
            "gbs": round(percall_gbs, 2),
            "encode_gbs": round(enc_bytes / t_enc / 1e9, 2),
            "reconstruct_gbs": round(rec_bytes / t_rec / 1e9, 2),
            "reconstruct_latency_us": {"p50": round(float(np.percentile(lat, 50)), 1),
                                       "p99": round(float(np.percentile(lat, 99)), 1)},
        },
        "batched_vs_per_call": round(value / percall_gbs, 2),
        "c1": c1,
        "host_path": host,
        "wide": wide,
        "tower": tower,
        "cost_model": costmodel,
        "served": served,
        "byte_exact": {"batched_reconstruct": rec_exact, "batched_encode": enc_exact,
                       "after_timed_steps": exact_after, "per_call": exact_percall,
                       "oracle_sampled_parity": host_ok},
    }
    print(json.dumps(out))
    ok = rec_exact and enc_exact and exact_after and exact_percall and host_ok
    if tower is not None:
        ok = ok and all(tower["byte_exact"].values())
    if c1 is not None:
        ok = ok and all(c1["byte_exact"].values())
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
