"""Pins of the MLP-preconditioner oracle (CPU only): BF16 rounding against
torch, backward against finite differences, the pair-loss gradient, the
correction terms against a dense solve, and the zero-output-layer network
giving the damped Neumann correction z - w/2."""
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
    go = rng.standard_normal((7, M.OUT))
    g = M.backward(p, X, go, exact=True)
    f = lambda q: float(np.sum(go * M.forward(q, X, exact=True)[0]))
    idx = np.concatenate([rng.choice(np.arange(M.W1, M.B1), 12, replace=False),
                          rng.choice(np.arange(M.B1, M.W2), 6, replace=False),
                          rng.choice(np.arange(M.W2, M.B2), 12, replace=False),
                          rng.choice(np.arange(M.B2, M.W3), 6, replace=False),
                          rng.choice(np.arange(M.W3, M.B3), 8, replace=False), [M.B3, M.B3 + 1]])
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
    go = rng.standard_normal((300, M.OUT)).astype(np.float32)
    ge = M.backward(p, M.bf16(X), go, exact=True)
    gq = M.backward(p, X, go)
    # BF16 weights and activations: a few per cent, not more
    assert np.linalg.norm(gq - ge) / np.linalg.norm(ge) < 0.1


def test_zero_output_layer_is_damped_neumann():
    p = M.init_params(9)
    p[M.W3:] = 0.0
    rng = np.random.default_rng(4)
    X = rng.standard_normal((50, M.IN)).astype(np.float32)
    o, _ = M.forward(p, X)
    assert o.shape == (50, 2) and np.all(o == 0.0)
    z = rng.standard_normal(50) + 1j * rng.standard_normal(50)
    w = rng.standard_normal(50) + 1j * rng.standard_normal(50)
    w2 = rng.standard_normal(50) + 1j * rng.standard_normal(50)
    assert np.array_equal(M.combine(o, z, w, w2), z - 0.5 * w)
    # the coefficients stay inside their bounds for any output
    s1, s2 = M.coefficients(np.array([[1e30, -1e30], [-1e30, 1e30]]))
    assert np.allclose(s1, [0.0, -1.0]) and np.allclose(s2, [-0.5, 0.5])


def test_neumann_terms_against_dense_solve():
    """z - w and z - w + w2 are the first- and second-order expansions of
    (H - theta)^-1 r in the off-diagonal part: scaling the off-diagonal by
    eps leaves errors of order eps^2 and eps^3 (a sign or index slip in w
    or w2 leaves a lower order)."""
    rng = np.random.default_rng(8)
    n = 30
    D = np.diag(rng.uniform(1.0, 3.0, n))
    O = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    O = O + O.conj().T
    np.fill_diagonal(O, 0.0)
    theta = -1.0
    r = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    errs, errs2 = [], []
    for eps in (1e-2, 1e-3):
        H = D + eps * O
        z, w, w2 = M.neumann_terms(np.diag(H).real, theta, r, lambda x: H @ x)
        exact = np.linalg.solve(H - theta * np.eye(n), r)
        errs.append(np.linalg.norm(z - w - exact) / np.linalg.norm(exact))
        errs2.append(np.linalg.norm(z - w + w2 - exact) / np.linalg.norm(exact))
        # Jacobi alone is first order
        assert np.linalg.norm(z - exact) / np.linalg.norm(exact) > 10 * errs[-1]
        assert errs[-1] > 10 * errs2[-1]
    assert errs[1] < errs[0] / 50.0, errs
    assert errs2[1] < errs2[0] / 500.0, errs2


def test_pair_loss_gradient_matches_finite_differences():
    rng = np.random.default_rng(6)
    n = 40
    c = lambda: rng.standard_normal(n) + 1j * rng.standard_normal(n)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A = A + A.conj().T
    H = lambda x: A @ x
    z, w, w2, u, b = c(), c(), c(), c(), c()
    u /= np.linalg.norm(u)
    theta = -2.5
    o = rng.standard_normal((n, 2)) * 0.5
    L, g = M.pair_loss(o, z, w, w2, b, u, theta, H)
    assert 0.0 <= L <= 1.0 and g.shape == (n, 2)
    h = 1e-7
    for i in range(n):
        for k in range(2):
            e = np.zeros((n, 2))
            e[i, k] = h
            fd = (M.pair_loss(o + e, z, w, w2, b, u, theta, H)[0]
                  - M.pair_loss(o - e, z, w, w2, b, u, theta, H)[0]) / (2 * h)
            assert abs(fd - g[i, k]) < 1e-6 * max(1.0, abs(fd)), (i, k, fd, g[i, k])
    P = lambda x: x - u * np.vdot(u, x)
    K = lambda x: P(A @ x - theta * x)
    # the loss vanishes when t is the accurate correction up to scale (and a
    # component along u, which the projection removes) ...
    t = P(M.combine(o, z, w, w2))
    L0, _ = M.pair_loss(o, z, w, w2, K(3.0 * t), u, theta, H)
    assert L0 < 1e-12
    # ... and is 1 when K t is orthogonal to K t*
    q = K(t)
    b1 = c()
    b1 -= q * np.vdot(q, b1) / np.vdot(q, q)
    L1, _ = M.pair_loss(o, z, w, w2, b1, u, theta, H)
    assert abs(L1 - 1.0) < 1e-12
    # the accurate correction of a Ritz pair gives b = -r (K t* = -r)
    Q, _ = np.linalg.qr(rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3)))
    ev, S = np.linalg.eigh(Q.conj().T @ A @ Q)
    uu, th = Q @ S[:, 0], ev[0]
    r = A @ uu - th * uu
    Pm = np.eye(n) - np.outer(uu, uu.conj())
    ts = np.linalg.solve(Pm @ (A - th * np.eye(n)) @ Pm + np.outer(uu, uu.conj()), -r)
    Kt = Pm @ (A @ ts - th * ts)
    assert np.linalg.norm(Kt + r) < 1e-10 * np.linalg.norm(r)


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
