"""The float64 MLP oracle's gradients against central finite differences."""
import numpy as np

from oracle import mlp


def test_gradients_finite_differences():
    rng = np.random.default_rng(0)
    sizes = [5, 7, 6, 9]   # 8 bins + 1 log-variance
    params = [(rng.normal(size=(o, i)) * 0.5, rng.normal(size=o) * 0.1) for i, o in zip(sizes[:-1], sizes[1:])]
    x = rng.normal(size=(4, 5))
    t = rng.random((4, 8))
    t /= t.sum(axis=1, keepdims=True)
    tv = rng.normal(size=4)
    loss, grads = mlp.loss_and_grad(params, x, t, tv, 0.3, 0.7)
    eps = 1e-6
    for li, (W, b) in enumerate(params):
        for arr, g in ((W, grads[li][0]), (b, grads[li][1])):
            for idx in [tuple(rng.integers(0, s) for s in arr.shape) for _ in range(5)]:
                old = arr[idx]
                arr[idx] = old + eps
                lp, _ = mlp.loss_and_grad(params, x, t, tv, 0.3, 0.7)
                arr[idx] = old - eps
                lm, _ = mlp.loss_and_grad(params, x, t, tv, 0.3, 0.7)
                arr[idx] = old
                assert abs((lp - lm) / (2 * eps) - g[idx]) < 1e-6 * max(1, abs(g[idx]))


def test_softmax_ce_minimum_at_target():
    """The head loss is minimised (gradient zero) when softmax(logits) = target."""
    t = np.array([[0.1, 0.2, 0.7]])
    W = np.eye(4)[:, :4]
    params = [(np.eye(4), np.zeros(4))]
    x = np.concatenate([np.log(t), [[0.5]]], axis=1)
    _, grads = mlp.loss_and_grad(params, x, t, np.array([0.5]), 1.0)
    assert np.abs(grads[0][1]).max() < 1e-12
