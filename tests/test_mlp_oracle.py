"""CPU pin of the float64 MLP reference (oracle/mlp_ref.py): its analytic
gradient equals central finite differences of its loss."""
import numpy as np

import mlp_ref as M


def test_reference_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = rng.normal(0, 0.3, M.COUNT)
    X = rng.normal(0, 1, (40, M.IN))
    T = rng.normal(0, 1, 40)
    _, g = M.gradients(p, X, T)
    idx = rng.choice(M.COUNT, 60, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros(M.COUNT)
        e[i] = h
        fd = (M.loss(p + e, X, T) - M.loss(p - e, X, T)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])
