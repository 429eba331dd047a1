"""CPU checks of the warm start network's oracle (oracle/mlp.py): its
float64 back-propagation against central finite differences, and its bf16
rounding against values the format fixes."""
import numpy as np
import pytest

from oracle import mlp


def test_bf16_rounding_fixed_points():
    # representable values stay; halfway cases go to even; the mantissa is 8 bits
    x = np.array([1.0, -2.5, 1 + 2.0 ** -7, 1 + 2.0 ** -8, 1 + 3 * 2.0 ** -8, 1 + 2.0 ** -8 + 2.0 ** -12, 3.0e38],
                 np.float32)
    want = np.array([1.0, -2.5, 1 + 2.0 ** -7, 1.0, 1 + 2.0 ** -6, 1 + 2.0 ** -7, 2.9979e38])
    got = mlp.bf16(x)
    assert np.array_equal(got[:6], want[:6])
    assert abs(got[6] - 3.0e38) / 3.0e38 < 2.0 ** -8


@pytest.mark.parametrize("seed", [0, 1])
def test_gradients_match_finite_differences(seed):
    rng = np.random.default_rng(seed)
    n_in, h, n_out, B = 5, 8, 3, 7
    count = n_in * h + h + h * h + h + h * n_out + n_out
    theta = rng.normal(0, 0.5, count)
    x = rng.normal(0, 1, (B, n_in))
    t = rng.normal(0, 1, (B, n_out))
    loss, g = mlp.gradients(theta, x, t, n_in, h, n_out, use_bf16=False)
    eps = 1e-6
    fd = np.zeros(count)
    for k in range(count):
        tp, tm = theta.copy(), theta.copy()
        tp[k] += eps
        tm[k] -= eps
        lp, _ = mlp.gradients(tp, x, t, n_in, h, n_out, use_bf16=False)
        lm, _ = mlp.gradients(tm, x, t, n_in, h, n_out, use_bf16=False)
        fd[k] = (lp - lm) / (2 * eps)
    assert np.max(np.abs(g - fd)) <= 1e-6 * max(1.0, np.max(np.abs(fd)))
    # the loss is the mean squared error of the forward pass
    y, _, _ = mlp.forward(theta, x, n_in, h, n_out, use_bf16=False)
    assert abs(loss - np.mean((y - t) ** 2)) <= 1e-14


def test_planted_gradient_mistake_is_caught():
    """A dropped ReLU mask in the backward pass (a plausible mistake) breaks
    the finite-difference agreement: the check above has teeth."""
    rng = np.random.default_rng(3)
    n_in, h, n_out, B = 4, 8, 2, 6
    count = n_in * h + h + h * h + h + h * n_out + n_out
    theta = rng.normal(0, 0.7, count)
    x, t = rng.normal(0, 1, (B, n_in)), rng.normal(0, 1, (B, n_out))
    layers = mlp.unpack(theta, n_in, h, n_out)
    y, zs, ins = mlp.forward(theta, x, n_in, h, n_out, use_bf16=False)
    dz = 2.0 * (y - t) / (B * n_out)
    W3 = layers[2][0]
    da = dz @ W3.T                        # no mask applied: the mistake
    gW2_wrong = ins[1].T @ da
    _, g = mlp.gradients(theta, x, t, n_in, h, n_out, use_bf16=False)
    off = n_in * h + h
    gW2 = g[off:off + h * h].reshape(h, h)
    assert np.max(np.abs(gW2_wrong - gW2)) > 1e-3 * np.max(np.abs(gW2))
