"""Pins the MLP oracle (oracle/mlp_ref.py) before it checks the GPU kernels: its float64
backward pass equals central finite differences of its loss, and its bf16 rounding is
round-to-nearest-even (checked on hand-picked bit patterns)."""
import numpy as np

from oracle import mlp_ref


def small_net(rng, sizes=(12, 9, 8, 7, 5)):
    return [(rng.normal(0, 0.5, (sizes[i + 1], sizes[i])), rng.normal(0, 0.1, sizes[i + 1])) for i in range(4)]


def test_backward_equals_finite_differences():
    rng = np.random.default_rng(0)
    params = small_net(rng)
    x = rng.normal(size=(6, 12))
    y = rng.integers(0, 5, 6)
    _, grads = mlp_ref.loss_and_grads(x, params, y, bf16=False)
    eps = 1e-6
    for layer in range(4):
        for which in (0, 1):
            p = params[layer][which]
            flat = p.reshape(-1)
            for idx in rng.choice(flat.size, min(flat.size, 6), replace=False):
                old = flat[idx]
                flat[idx] = old + eps
                lp, _ = mlp_ref.loss_and_grads(x, params, y, bf16=False)
                flat[idx] = old - eps
                lm, _ = mlp_ref.loss_and_grads(x, params, y, bf16=False)
                flat[idx] = old
                fd = (lp - lm) / (2 * eps)
                g = grads[layer][which].reshape(-1)[idx]
                assert abs(fd - g) <= 1e-6 + 1e-5 * abs(fd), (layer, which, idx, fd, g)


def test_loss_is_log_classes_for_zero_logits():
    params = [(np.zeros((4, 3)), np.zeros(4)), (np.zeros((4, 4)), np.zeros(4)), (np.zeros((4, 4)), np.zeros(4)),
              (np.zeros((10, 4)), np.zeros(10))]
    loss, _ = mlp_ref.loss_and_grads(np.ones((3, 3)), params, np.array([0, 4, 9]), bf16=False)
    assert abs(loss - np.log(10)) < 1e-12


def test_bf16_rounding():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.0e38], np.float32)
    want = np.array([1.0, 1.0, 1.0 + 2 ** -6, 1.0, -2.5, 2.9983e38])
    got = mlp_ref.bf16_round(vals)
    assert np.array_equal(got[:5], want[:5])
    assert abs(got[5] - 3.0e38) / 3.0e38 < 2 ** -8
    assert np.array_equal(mlp_ref.bf16_bits_to_f64(np.array([0x3F80, 0xC020], np.uint16)), [1.0, -2.5])
