"""GPU checks of the learned cost model and the autotuner.

* The HIP MLP's forward pass and gradients equal the float64 CPU reference
  (oracle/mlp_ref.py) within fp32 tolerance (rel 2e-3 on gradients, abs 1e-4
  on outputs), and training lowers the loss.
* trs_tune_apply times every feasible kernel variant and rejects infeasible
  ones; candidates, features and rules come from the library.
* On held-out shapes the trained model's measured top 3 is within 10 % of the
  exhaustive optimum on average and beats both methods without the model.
* Installed in the library, the model routes real codec calls (byte-exact).
"""
import numpy as np
import pytest

import mlp_ref as M

pytestmark = pytest.mark.gpu


def _model(trs):
    from reed_solomon_amd import autotune as A

    return A, A.CostModel(seed=7)


def test_forward_and_gradients_match_cpu_reference(trs):
    A, cm = _model(trs)
    rng = np.random.default_rng(3)
    X = rng.normal(0, 1, (300, M.IN)).astype(np.float32)
    T = rng.normal(0, 1, 300).astype(np.float32)
    p = cm.params().astype(np.float64)
    y_ref, _ = M.forward(p, X)
    y = cm.raw_predict(X)
    assert np.allclose(y, y_ref, atol=1e-4, rtol=1e-4)
    loss, g = cm.raw_gradients(X, T)
    loss_ref, g_ref = M.gradients(p, X, T)
    assert abs(loss - loss_ref) <= 1e-4 * max(1.0, abs(loss_ref))
    scale = np.abs(g_ref).max()
    assert np.abs(g - g_ref).max() <= 2e-3 * scale


def test_training_lowers_loss(trs):
    A, cm = _model(trs)
    rng = np.random.default_rng(4)
    X = rng.normal(0, 1, (2000, M.IN)).astype(np.float32)
    T = (np.sin(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2]).astype(np.float32)
    before, _ = cm.raw_gradients(X, T)
    after = cm.raw_train(X, T, 400, 3e-3)
    assert after < 0.5 * before


def test_tune_variants(trs):
    from reed_solomon_amd import autotune as A

    V = {n: i for i, n in enumerate(A.VARIANTS)}
    # a generated GF(2^8) shape: matrix kernels, MFMA and its bit-sliced program
    s = A.Shape(8, 10, 10, 1 << 16)
    got = A.exhaustive(s)
    assert set(got) == {V["table"], V["planes4"], V["planes8"], V["mfma"], V["program32k"]}
    assert all(t > 0 for t in got.values())
    assert A.rule(s) == V["program32k"]
    # at most 4 outputs: no separate 8-row planes candidate
    assert V["planes8"] not in A.candidates(A.Shape(8, 12, 4, 1 << 20))
    # tower shapes: the matrix kernel and the FFT program's distinct LDS targets
    s16 = A.Shape(16, 40, 12, 8192)
    c16 = A.candidates(s16)
    assert V["table"] in c16 and any(v >= V["program16k"] for v in c16)
    assert all(A.measure(s16, v) > 0 for v in c16)
    assert A.measure(A.Shape(16, 8, 4100, 4096), V["program32k"]) is None
    assert A.measure(A.Shape(8, 40, 5, 4096), V["mfma"]) is None
    # GF(2^8) shapes without a generated program: the staged program, any length
    s8 = A.Shape(8, 64, 64, 1 << 20)
    assert V["mfma"] not in A.candidates(s8)
    assert any(v >= V["program16k"] for v in A.candidates(s8))
    assert A.measure(A.Shape(8, 64, 64, 4097), max(A.candidates(A.Shape(8, 64, 64, 4097)))) > 0
    assert A.heuristic(s8) >= V["program16k"]
    assert A.heuristic(A.Shape(8, 12, 8, 1 << 20)) < V["program16k"]
    # the model's inputs come from the library and carry the variant's kind
    f = A.features(s, V["mfma"])
    assert f.shape == (12,) and f[8] == 1.0 and f[6] == f[7] == f[9] == 0.0


def _study(A, n_train, n_test, seed):
    shapes = A.sample_shapes(n_train + n_test, seed)
    timings = [A.exhaustive(s) for s in shapes]
    X, T = A.dataset(shapes[:n_train], timings[:n_train])
    cm = A.CostModel(seed=1)
    cm.fit(X, T, steps=1500)
    return shapes[n_train:], timings[n_train:], cm


def test_model_regret_beats_the_rules_on_held_out_shapes(trs):
    """The trained network's measured top 3 stays within 10 % of the
    exhaustive optimum on held-out shapes, and beats both methods without it:
    the codec's own GPU-calibrated rules and the reference's routing
    (src/fft16.rs:85-106, src/reedsolomon.rs:204-234)."""
    from reed_solomon_amd import autotune as A

    held, meas, cm = _study(A, 200, 60, 0xC0DE)
    tuner = A.Tuner(cm, top_n=3)
    model = np.array([A.regret(t, tuner.choose(s, t)) for s, t in zip(held, meas)])
    rules = np.array([A.regret(t, A.rule(s)) for s, t in zip(held, meas)])
    ref = np.array([A.regret(t, A.heuristic(s)) for s, t in zip(held, meas)])
    print(f"mean regret: model top-3 {model.mean():.4f}, codec rules {rules.mean():.4f}, "
          f"reference routing {ref.mean():.4f}")
    assert model.mean() <= 0.10
    assert model.mean() < rules.mean()
    assert model.mean() < ref.mean()
    # measuring only the top 3 of the ranked candidates is fewer runs than
    # measuring them all
    assert np.mean([min(3, len(t)) for t in meas]) < np.mean([len(t) for t in meas])


def test_router_in_the_codec_path(trs):
    """With the trained model installed, automatic encodes and matrix-plan
    reconstructs rank, time and cache a kernel variant per key; outputs stay
    byte-identical to the oracle, and removing the model restores the rules."""
    import torch

    import oracle as O
    from reed_solomon_amd import autotune as A
    from util import rows_to_dev, seeded, to_host

    _, _, cm = _study(A, 80, 0, 0xBEEF)
    cm.install(top_n=3)
    try:
        for (k, m, length) in ((20, 10, 300_000), (64, 64, 65_536), (7, 13, 100_000)):
            rs = trs.ReedSolomon(k, m)
            G = O.build_matrix(8, k, k + m)
            data = seeded(k, length, 17 + k)
            for rep in range(2):
                shards = rows_to_dev(np.concatenate([data, np.zeros((m, length), np.uint8)]))
                rs.encode(shards)
                torch.cuda.synchronize()
                assert (to_host(shards[k:]) == O.rs8_encode(G, k, data)).all(), (k, m, rep)
                assert A.routed_choice(8, 0, k, m, length) in A.candidates(A.Shape(8, k, m, length))
            # a matrix-plan reconstruct of the same stripe
            full = np.concatenate([data, O.rs8_encode(G, k, data)])
            erased = list(range(0, k + m, 3))[:m]
            cur = full.copy()
            cur[erased] = 0
            rows = rows_to_dev(cur)
            rs.set_decode_route("matrix")
            rs.reconstruct([(t, i not in erased) for i, t in enumerate(rows)])
            torch.cuda.synchronize()
            assert (to_host(rows) == full).all()
            nout = len(erased)
            assert A.routed_choice(8, 1, k, nout, length) >= 0
        rs16 = trs.ReedSolomon16(100, 28)
        data = seeded(100, 8192, 3)
        shards = rows_to_dev(np.concatenate([data, np.zeros((28, 8192), np.uint8)]))
        rs16.encode(shards)
        torch.cuda.synchronize()
        want = O.rs16_symbolwise(O.build_matrix(16, 100, 128)[100:], list(data), 8192)
        assert (to_host(shards[100:]) == want).all()
        assert A.routed_choice(16, 0, 100, 28, 8192) >= 0
    finally:
        A.uninstall()
    assert A.routed_choice(8, 0, 20, 10, 300_000) == -1
