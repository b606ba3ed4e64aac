"""GPU checks of the learned cost model and the autotuner.

* The HIP MLP's forward pass and gradients equal the float64 CPU reference
  (oracle/mlp_ref.py) within fp32 tolerance (rel 2e-3 on gradients, abs 1e-4
  on outputs), and training lowers the loss.
* trs_tune_apply times every feasible route and rejects infeasible ones.
* A small autotune study runs end to end: the model-guided choice's regret is
  a finite, non-negative number for every held-out shape.
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


def test_tune_apply_routes(trs):
    from reed_solomon_amd import autotune as A

    s = A.Shape(8, 10, 10, 1 << 16)
    got = A.exhaustive(s)
    routes = {c[0] for c in got}
    assert routes == {"table", "bitsliced", "mfma"}
    assert all(t > 0 for t in got.values())
    s16 = A.Shape(16, 40, 12, 8192)
    assert {c[0] for c in A.exhaustive(s16)} == {"table", "bitsliced"}
    # the tower FFT candidate over strips equals nothing in output terms, but
    # must time every strip size; shapes past k + m = 4096 have none
    assert A.measure(A.Shape(16, 1000, 24, 65536), ("bitsliced", 8192)) > 0
    assert A.measure(A.Shape(16, 8, 4100, 4096), ("bitsliced", 0)) is None
    assert A.measure(A.Shape(8, 40, 5, 4096), ("mfma", 0)) is None
    # GF(2^8) shapes without a generated program: the staged FFT program on
    # the byte-plane interpreter, any length and strip
    s8 = A.Shape(8, 64, 64, 1 << 20)
    assert {c[0] for c in A.candidates(s8)} == {"table", "bitsliced"}
    assert A.measure(s8, ("bitsliced", 65536)) > 0
    assert A.measure(A.Shape(8, 64, 64, 4097), ("bitsliced", 1000)) > 0
    assert A.heuristic(s8) == ("bitsliced", 0)
    assert A.heuristic(A.Shape(8, 12, 8, 1 << 20)) == ("table", 0)


def test_small_study_end_to_end(trs):
    from reed_solomon_amd import autotune as A

    shapes = A.sample_shapes(60, 11)
    timings = [A.exhaustive(s, cap_bytes=4 << 20) for s in shapes]
    X, T = A.dataset(shapes[:45], timings[:45])
    cm = A.CostModel(seed=1)
    cm.fit(X, T, steps=300)
    tuner = A.Tuner(cm, top_n=3)
    for s, meas in zip(shapes[45:], timings[45:]):
        r = A.regret(meas, tuner.choose(s, meas))
        assert 0.0 <= r < 100.0
