"""The device MLP (matrix-core products) against the CPU oracle: forward
values, loss and every gradient entry; and training on the device lowers
the loss on a learnable map."""
import numpy as np
import pytest

from eprem_amd import runtime
from oracle import mlp

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("sizes,n", [([96, 128, 128, 128, 30], 200), ([10, 33, 17, 5], 75)])
def test_forward_and_gradient_match_oracle(sizes, n):
    rng = np.random.default_rng(3)
    net = runtime.Mlp(sizes, seed=11)
    try:
        p = net.params()
        x = rng.normal(0, 1, (n, sizes[0])).astype(np.float32)
        y = rng.normal(0, 1, (n, sizes[-1])).astype(np.float32)
        out = net.forward(x)
        loss, g = net.loss_grad(x, y)
    finally:
        net.close()
    acts = mlp.forward(p.astype(np.float64), sizes, x)
    np.testing.assert_allclose(out, acts[-1], rtol=1e-4, atol=1e-5)
    lo, go = mlp.loss_grad(p.astype(np.float64), sizes, x, y)
    assert loss == pytest.approx(lo, rel=1e-4)
    scale = np.abs(go).max()
    np.testing.assert_allclose(g, go, rtol=2e-3, atol=2e-4 * scale)


@pytest.mark.parametrize("sizes,n", [([96, 128, 128, 128, 30], 1000), ([66, 128, 128, 20], 333),
                                     ([128, 128, 32], 64), ([7, 128, 128, 128, 128, 1], 129)])
def test_fused_forward_matches_oracle(sizes, n):
    """Networks of the surrogate's shape (input <= 128, hidden layers of
    128, output <= 32) run in the one-kernel forward (weights in LDS,
    activations in registers); inputs that are not a multiple of 16 wide,
    row counts that are not a multiple of the 64-row wave tile, one to four
    hidden layers."""
    rng = np.random.default_rng(5)
    net = runtime.Mlp(sizes, seed=13)
    try:
        p = net.params()
        p[np.arange(p.size) % 7 == 3] *= -1.5  # biases and weights of either sign
        net.set_params(p)
        x = rng.normal(0, 1, (n, sizes[0])).astype(np.float32)
        out = net.forward(x)
    finally:
        net.close()
    ref = mlp.forward(p.astype(np.float64), sizes, x)[-1]
    # Activations are rounded to bf16 between layers; the device accumulates
    # in fp32 (in another order than the oracle's float64), so an activation
    # within an fp32 rounding of a bf16 tie can round to the neighbouring
    # bf16 value (2^-8 relative) and move a few outputs by up to ~1e-3 of
    # the output scale.  Bar: 99.5% of outputs within 1e-4, all within 1e-2.
    scale = np.abs(ref).max()
    err = np.abs(out - ref)
    tight = err <= 1e-4 * np.abs(ref) + 1e-4 * scale
    assert tight.mean() >= 0.995, tight.mean()
    assert err.max() <= 1e-2 * scale, err.max() / scale


def test_training_lowers_loss():
    rng = np.random.default_rng(4)
    sizes = [8, 64, 64, 4]
    x = rng.normal(0, 1, (4096, 8)).astype(np.float32)
    A = rng.normal(0, 0.5, (8, 4))
    y = (np.tanh(x @ A) + 0.1 * x[:, :4] ** 2).astype(np.float32)
    net = runtime.Mlp(sizes, seed=5)
    try:
        hist = net.train(x, y, epochs=30, batch=256, lr=3e-3)
        pred = net.forward(x)
    finally:
        net.close()
    assert hist[-1] < 0.1 * hist[0], hist
    assert np.mean((pred - y) ** 2) < 0.2 * np.mean((y - y.mean(0)) ** 2)
