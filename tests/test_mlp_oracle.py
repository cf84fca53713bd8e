"""Pins the surrogate-MLP oracle: its backward pass against central finite differences of its own
forward pass in float64 (no rounding), and its bfloat16 rounding against the textbook rule."""
import numpy as np

import mlp_oracle as mo


def test_bf16_rounding_is_nearest_even():
    # bfloat16 keeps 7 fraction bits: the neighbours of 1 are 1 - 2^-8 and 1 + 2^-7
    assert mo.bf16(np.float32(1.0)) == 1.0
    assert mo.bf16(np.float32(1.0 + 2 ** -9)) == 1.0               # below half an ulp: down
    assert mo.bf16(np.float32(1.0 + 3 * 2 ** -9)) == 1.0 + 2 ** -7  # above half an ulp: up
    # ties go to even: 1 + 2^-8 is halfway between 1 and 1 + 2^-7 -> 1
    assert mo.bf16(np.float32(1.0 + 2 ** -8)) == 1.0
    assert mo.bf16(np.float32(1.0 + 3 * 2 ** -8)) == np.float32(1.0 + 2 ** -6)
    assert mo.bf16(np.float32(-2.5)) == -2.5
    # every result is a bfloat16 (low 16 bits zero) within half an ulp
    v = np.random.default_rng(0).standard_normal(10000).astype(np.float32)
    b = mo.bf16(v)
    assert np.all((b.view(np.uint32) & 0xFFFF) == 0)
    assert np.all(np.abs(b - v) <= np.abs(v) * 2 ** -8 + 1e-45)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    F, W, D = 8, 16, 3
    n = 37
    Ks = [2 * F] + [W] * (D - 1)
    npar = sum(K * W for K in Ks) + D * W + W + 1
    p = rng.standard_normal(npar) * 0.3
    B = rng.standard_normal((F, 3)) * 2.0
    x = rng.uniform(-1, 1, (n, 3))
    y = rng.standard_normal(n) * 0.1
    loss, g = mo.loss_grad(p, B, x, y, F, W, D, emulate=False)
    idx = rng.choice(npar, 60, replace=False)
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        lp, _ = mo.loss_grad(pp, B, x, y, F, W, D, emulate=False)
        lm, _ = mo.loss_grad(pm, B, x, y, F, W, D, emulate=False)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
