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


def test_forward_on_a_hand_built_network():
    """Every weight zero but one path: x -> feature sin(2 pi B_0 . x) -> unit 0 of each layer ->
    output, so phi = w_out0 silu(... silu(f + b_0) ...) + b_out is known in closed form."""
    F, W, D = 8, 16, 3
    Ks = [2 * F] + [W] * (D - 1)
    npar = sum(K * W for K in Ks) + D * W + W + 1
    p = np.zeros(npar)
    B = np.zeros((F, 3))
    B[0] = (0.5, -0.25, 1.0)
    off = 0
    p[off + 0 * W + 0] = 1.5           # W_0[feature 0 (sin of frequency 0), unit 0]
    off += Ks[0] * W
    for l in range(1, D):
        p[off + 0 * W + 0] = 0.75      # W_l[unit 0, unit 0]
        off += W * W
    for l in range(D):
        p[off + l * W + 0] = 0.1 * (l + 1)  # b_l[unit 0]
    off += D * W
    p[off + 0] = 2.0                   # w_out[unit 0]
    p[off + W] = -0.3                  # b_out
    x = np.array([[0.2, 0.4, -0.1], [-0.7, 0.3, 0.9]])
    silu = lambda z: z / (1 + np.exp(-z))
    want = []
    for q in x:
        a = silu(1.5 * np.sin(2 * np.pi * (0.5 * q[0] - 0.25 * q[1] + 1.0 * q[2])) + 0.1)
        for l in range(1, D):
            a = silu(0.75 * a + 0.1 * (l + 1))
        want.append(2.0 * a - 0.3)
    got, _ = mo.forward(p, B, x, F, W, D, emulate=False)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12)
