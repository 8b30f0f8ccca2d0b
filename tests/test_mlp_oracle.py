"""Pins of the NumPy warm-start network (CPU)."""

import numpy as np

from oracle import mlp_np as N


def test_bf16_round_matches_bit_definition():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 65504.0, 1e-30, 0.0], np.float32)
    got = N.bf16_round(x)
    # truncate-with-rounding by hand: keep the top 16 bits, ties to even
    for a, g in zip(x, got):
        u = int(np.float32(a).view(np.uint32))
        low = u & 0xFFFF
        hi = u >> 16
        if low > 0x8000 or (low == 0x8000 and (hi & 1)):
            hi += 1
        ref = np.uint32(hi << 16).view(np.float32)
        assert g == np.float64(ref)
    # exact tie 1 + 2^-8 rounds to even (1.0), 1 + 3*2^-9 rounds up
    assert N.bf16_round(np.float32(1.0 + 2 ** -8)) == 1.0
    assert N.bf16_round(np.float32(1.0 + 3 * 2 ** -9)) == 1.0 + 2 ** -7


def test_gelu_grad_matches_finite_differences():
    z = np.linspace(-4, 4, 41)
    h = 1e-6
    fd = (N.gelu(z + h) - N.gelu(z - h)) / (2 * h)
    np.testing.assert_allclose(N.gelu_grad(z), fd, atol=1e-8)


def test_mlp_gradient_matches_finite_differences():
    shapes = N.layer_shapes(6, 8, 2, 3)
    rng = np.random.default_rng(0)
    p = N.init_params(6, 8, 2, 3, seed=1).astype(np.float64)
    p += 0.05 * rng.standard_normal(p.size)      # non-zero biases too
    X = rng.standard_normal((5, 6))
    T = rng.standard_normal((5, 3))
    M = (rng.random((5, 3)) > 0.3).astype(float)
    loss, g = N.loss_and_grad(p, X, T, M, shapes, bf16=False)
    h = 1e-6
    fd = np.empty_like(p)
    for i in range(p.size):
        e = np.zeros_like(p)
        e[i] = h
        fd[i] = (N.loss_and_grad(p + e, X, T, M, shapes, bf16=False)[0]
                 - N.loss_and_grad(p - e, X, T, M, shapes, bf16=False)[0]) / (2 * h)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-8)


def test_adam_first_step_moves_by_lr():
    p = np.array([1.0, -2.0, 0.5], np.float32)
    g = np.array([0.3, -0.1, 0.0], np.float32)
    m, v = np.zeros(3, np.float32), np.zeros(3, np.float32)
    N.adam_step(p, g, m, v, 1, lr=0.01)
    # first bias-corrected step is lr * sign(g) (up to eps)
    np.testing.assert_allclose(p, [0.99, -1.99, 0.5], atol=1e-6)


def test_adamw_steps_match_torch():
    """Several bias-corrected AdamW steps with weight decay (t = 1..6) against
    torch.optim.AdamW on the CPU (decoupled decay: p <- p - lr (m^/(sqrt v^ + eps)
    + wd p))."""
    import torch
    rng = np.random.default_rng(4)
    p0 = rng.standard_normal(50).astype(np.float32)
    grads = [rng.standard_normal(50).astype(np.float32) * (0.1 + t) for t in range(6)]
    lr, b1, b2, eps, wd = 3e-3, 0.9, 0.999, 1e-8, 1e-2
    p = p0.copy()
    m, v = np.zeros(50, np.float32), np.zeros(50, np.float32)
    tp = torch.nn.Parameter(torch.tensor(p0.copy()))
    opt = torch.optim.AdamW([tp], lr=lr, betas=(b1, b2), eps=eps, weight_decay=wd)
    for t, g in enumerate(grads, start=1):
        N.adam_step(p, g, m, v, t, lr=lr, beta1=b1, beta2=b2, eps=eps, wd=wd)
        tp.grad = torch.tensor(g.copy())
        opt.step()
        np.testing.assert_allclose(p, tp.detach().numpy(), rtol=2e-6, atol=2e-7)
    assert np.abs(p - p0).max() > 3 * lr      # the steps moved the parameters
