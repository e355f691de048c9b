"""Pins of the MLP oracle: its gradient equals central finite differences
of its loss, and the bf16 rounding matches the definition of bf16."""
import numpy as np

from oracle import mlp


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    sizes = [5, 7, 6, 3]
    n = sum(a * b + b for a, b in zip(sizes[:-1], sizes[1:]))
    p = rng.normal(0, 0.5, n)
    x = rng.normal(0, 1, (11, 5))
    y = rng.normal(0, 1, (11, 3))
    _, g = mlp.loss_grad(p, sizes, x, y, exact=True)
    h = 1e-6
    for k in rng.choice(n, 25, replace=False):
        e = np.zeros(n)
        e[k] = h
        fd = (mlp.loss_grad(p + e, sizes, x, y, exact=True)[0] - mlp.loss_grad(p - e, sizes, x, y, exact=True)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])), (k, fd, g[k])


def test_bf16_rounding():
    # 1 + 2^-8 is exactly halfway between bf16 neighbours 1 and 1 + 2^-7:
    # ties go to even (1.0); 1 + 3*2^-9 rounds up
    v = mlp.to_bf16(np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, -2.5, 3.0e38], np.float32))
    assert v[0] == 1.0 and v[1] == 1.0 and v[2] == 1.0 + 2.0 ** -7 and v[3] == -2.5
    assert np.isfinite(v[4])
