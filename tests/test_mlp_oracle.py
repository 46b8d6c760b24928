"""Pins of the MLP-preconditioner oracle (CPU only): BF16 rounding against
torch, backward against finite differences, the pair-loss gradient, and the
zero-output-layer network reducing to the Jacobi correction."""
import numpy as np
import pytest
import torch

from oracle import mlp_oracle as M


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([
        rng.standard_normal(20000).astype(np.float32) * 10.0 ** rng.integers(-30, 30, 20000),
        np.array([0.0, -0.0, 1.0, 1.00390625, 1.01171875, 3.0e38, -3.0e38, 1e-40, -1e-42,
                  np.inf, -np.inf], dtype=np.float32),
    ]).astype(np.float32)
    # exact ties: mantissa ...x 1000 0000 0000 0000 in the discarded half
    ties = (np.arange(1, 2000, dtype=np.uint32) << 16 | 0x8000).view(np.float32)
    x = np.concatenate([x, ties, -ties])
    ours = M.bf16(x)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(ours.view(np.uint32), ref.view(np.uint32))
    assert np.isnan(M.bf16(np.array([np.nan], dtype=np.float32))[0])


def test_exact_backward_matches_finite_differences():
    p = M.init_params(3).astype(np.float64)
    rng = np.random.default_rng(1)
    X = rng.standard_normal((7, M.IN))
    go = rng.standard_normal(7)
    g = M.backward(p, X, go, exact=True)
    f = lambda q: float(go @ M.forward(q, X, exact=True)[0])
    idx = np.concatenate([rng.choice(np.arange(M.W1, M.B1), 12, replace=False),
                          rng.choice(np.arange(M.B1, M.W2), 6, replace=False),
                          rng.choice(np.arange(M.W2, M.B2), 12, replace=False),
                          rng.choice(np.arange(M.B2, M.W3), 6, replace=False),
                          rng.choice(np.arange(M.W3, M.B3), 6, replace=False), [M.B3]])
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (f(p + e) - f(p - e)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_rounded_backward_is_close_to_exact():
    p = M.init_params(5)
    rng = np.random.default_rng(2)
    X = rng.standard_normal((300, M.IN)).astype(np.float32)
    go = rng.standard_normal(300).astype(np.float32)
    ge = M.backward(p, M.bf16(X), go, exact=True)
    gq = M.backward(p, X, go)
    # BF16 weights and activations: a few per cent, not more
    assert np.linalg.norm(gq - ge) / np.linalg.norm(ge) < 0.1


def test_zero_output_layer_is_jacobi():
    p = M.init_params(9)
    p[M.W3:M.B3] = 0.0
    p[M.B3] = 0.0
    rng = np.random.default_rng(4)
    X = rng.standard_normal((50, M.IN)).astype(np.float32)
    o, _ = M.forward(p, X)
    assert np.all(o == 0.0)


def test_pair_loss_gradient_matches_finite_differences():
    rng = np.random.default_rng(6)
    n = 40
    diag = rng.uniform(-3, 1, n)
    theta = -4.0
    r = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    u = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    u /= np.linalg.norm(u)
    ts = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ts -= u * np.vdot(u, ts)
    o = rng.standard_normal(n) * 0.1
    L, g = M.pair_loss(diag, theta, o, r, ts, u)
    assert 0.0 <= L <= 1.0
    h = 1e-7
    for i in range(n):
        e = np.zeros(n)
        e[i] = h
        fd = (M.pair_loss(diag, theta, o + e, r, ts, u)[0] - M.pair_loss(diag, theta, o - e, r, ts, u)[0]) / (2 * h)
        assert abs(fd - g[i]) < 1e-6 * max(1.0, abs(fd))
    # the loss vanishes when the preconditioned residual is parallel to ts
    w = 1.0 / M.jacobi_den(diag, theta)
    o_star = (ts / r).real - w   # only exact when ts/r is real
    ts2 = (w + o_star) * r
    ts2 = ts2 - u * np.vdot(u, ts2)
    L2, _ = M.pair_loss(diag, theta, o_star, r, ts2, u)
    assert L2 < 1e-12


def test_features_definition():
    diag = np.array([-1.0, 0.5, 2.0, -4.0])
    occ = np.array([[0.5, 0.25, 0.25, 0.0]] * 4, dtype=np.float32)
    conn = np.array([0.1, 0.2, 0.3, 0.4], dtype=np.float32)
    f = M.features(diag, occ, conn, -4.0, 8)
    assert f.shape == (4, 32)
    assert np.all(f[:, 10:] == 0)
    # den = 0 clamps to +1e-10: 1/den clipped at 4 -> feature 1 = 1
    assert f[3, 1] == 1.0 and f[3, 0] == 0.0
    assert f[0, 2] == M.bf16(np.float32(-0.5))
