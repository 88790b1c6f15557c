"""CPU pins of the step-size network's oracle (oracle/mlp_oracle.py) and of
the feature extraction: the analytic gradient against central finite
differences, the parameter count, and the features against their
definitions on a hand-made history."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import mlp_oracle as M  # noqa: E402

from unicycle_amd import stepnet  # noqa: E402


def test_parameter_count():
    assert M.NP == 4801 == stepnet.N_PARAMS


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    th = rng.normal(0, 0.3, M.NP)
    X = rng.normal(0, 1, (40, 8))
    T = rng.normal(0, 1, 40)
    norm = (np.zeros(8), np.ones(8), 0.1, 2.0)
    loss, g = M.loss_grad(th, X, T, norm)
    idx = np.concatenate([rng.choice(M.NP, 60, replace=False), [0, M.OFF_B1, M.OFF_W2,
                                                                 M.OFF_B2, M.OFF_W3, M.OFF_B3]])
    h = 1e-6
    for i in idx:
        tp, tm = th.copy(), th.copy()
        tp[i] += h
        tm[i] -= h
        fd = (M.loss_grad(tp, X, T, norm)[0] - M.loss_grad(tm, X, T, norm)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), i


def test_features_from_history_definitions():
    dt = np.array([1e-3, 1e-2, 5e-2, 4e-2])
    vmax = np.array([1e-9, 2e-9, 1e-8, 1e-7])
    err = np.array([1e-5, 0.3, 0.8, 0.5])
    rej = np.array([0, 0, 2, 1])
    X, T = stepnet.features_from_history(dt, vmax, err, rej)
    assert X.shape == (3, 8) and T.shape == (3,)
    k = 2
    np.testing.assert_allclose(X[k], [np.log10(1e-8), np.log10(1e-8) - np.log10(2e-9),
                                      np.log10(5e-2), np.log10(0.8), np.log10(1e-2), 2,
                                      np.log10(0.7 * 5e-2 * 0.8 ** -0.2), np.log10(0.3)],
                               rtol=1e-6)
    # first step: trend 0, previous = current
    assert X[0, 1] == 0 and X[0, 4] == X[0, 2] and X[0, 7] == X[0, 3]
    # classic growth x10 below ERRCON
    assert np.isclose(X[0, 6], np.log10(1e-2))
    # target: next accepted step scaled to the target error ratio
    a = stepnet.ERR_TARGET
    np.testing.assert_allclose(T[k], np.log10(4e-2) + 0.2 * np.log10(a / 0.5), rtol=1e-6)
    np.testing.assert_allclose(T[0], np.log10(1e-2) + 0.2 * np.log10(a / 0.3), rtol=1e-6)
