"""Pins of the numpy oracle of the learned WCNS sensor: the analytic
gradient against central finite differences, and the features and labels
against cases whose answer is known."""
import numpy as np

from oracle import mlp


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    w = mlp.init_params(1).astype(np.float64)
    X = rng.random((64, mlp.NF))
    y = (rng.random(64) > 0.5).astype(np.float64)
    loss, g = mlp.loss_and_grad(w, X, y)
    h = 1e-6
    idx = rng.choice(mlp.NPARAM, 60, replace=False)
    for e in list(idx) + [mlp.NPARAM - 1, 0]:
        wp, wm = w.copy(), w.copy()
        wp[e] += h
        wm[e] -= h
        fd = (mlp.loss_and_grad(wp, X, y)[0] - mlp.loss_and_grad(wm, X, y)[0]) / (2 * h)
        assert abs(fd - g[e]) <= 1e-6 + 1e-5 * abs(g[e]), (e, fd, g[e])
    assert loss > 0


def test_features_are_scale_invariant_and_bounded():
    rng = np.random.default_rng(1)
    rho = 1 + 0.1 * rng.random((50, 6))
    p = 2 + 0.1 * rng.random((50, 6))
    f1 = mlp.features(rho, p)
    f2 = mlp.features(3.0 * rho, 3.0 * p)
    # the normalised indicators and jumps do not change with the scale
    for c in (0, 1, 2, 3, 4, 6, 7, 8, 9, 10):
        assert np.allclose(f1[:, c], f2[:, c], rtol=1e-12, atol=1e-14)
    assert np.allclose(f1[:, :4].sum(1), 1.0) and np.allclose(f1[:, 6:10].sum(1), 1.0)
    # the log-magnitude moves by log10(9)/10 per quantity
    assert np.allclose(f2[:, 5] - f1[:, 5], np.log10(9.0) / 10.0, atol=1e-9)


def test_labels_linear_data_and_a_jump():
    a = np.arange(6, dtype=np.float64)
    lin = np.stack([1.0 + 0.01 * a, 0.1 * a, 0 * a, 0 * a, 2.5 + 0.02 * a], -1)
    assert mlp.labels(lin) == 0.0  # equal indicators: linear weights exactly
    jump = lin.copy()
    jump[3:, 0] += 0.5
    assert mlp.labels(jump) == 1.0


def test_weights_reduce_to_linear_on_a_line():
    w = mlp.weno5_weights(1.0, 2.0, 3.0, 4.0, 5.0)
    assert np.allclose(w, mlp.DLIN, atol=1e-14)
