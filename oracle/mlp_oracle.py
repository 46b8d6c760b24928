"""CPU restatement of the learned Davidson preconditioner (test infrastructure).

Only tests/ may import this module; the library never calls it.  The learned
preconditioner has no counterpart in the reference, whose Davidson corrects
with the Jacobi diagonal (src/solvers/davidson.cpp:166-177): this file
restates the library's own definition in numpy, with the BF16 rounding at
the same points as the GPU kernels, so that the kernels' forward pass,
parameter gradients, features and pair loss can be checked.

Pinned by: float32->bfloat16 rounding against torch's conversion; the exact
(float64, unrounded) backward against central finite differences of the
exact forward; the pair-loss gradient against finite differences; the
network with a zero output layer equals the Jacobi method.

Network: x (32) -> relu(x W1 + b1) (128) -> relu(h1 W2 + b2) (128) -> h2 w3 + b3.
Parameter block (float32): W1[32][128], b1[128], W2[128][128], b2[128],
w3[128], b3, row index = input.
"""
from __future__ import annotations

import numpy as np

IN, HID = 32, 128
W1, B1 = 0, IN * HID
W2 = B1 + HID
B2 = W2 + HID * HID
W3 = B2 + HID
B3 = W3 + HID
NPARAMS = B3 + 1


def bf16(x) -> np.ndarray:
    """Round float32 values to the nearest bfloat16 (ties to even), returned
    as float32; NaN becomes the quiet NaN 0x7fc0."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, 0x7FC0, r).astype(np.uint32) << 16
    return r.view(np.float32).reshape(a.shape)


def unpack(p: np.ndarray):
    p = np.asarray(p)
    return (p[W1:B1].reshape(IN, HID), p[B1:W2], p[W2:B2].reshape(HID, HID), p[B2:W3],
            p[W3:B3], p[B3])


def init_params(seed: int) -> np.ndarray:
    """Any parameter block works with the checks; this one is not the library's
    initialisation (which uses std::mt19937 on the host)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAMS, dtype=np.float32)
    p[W1:B1] = rng.uniform(-1, 1, IN * HID) * np.sqrt(6.0 / IN)
    p[B1:W2] = rng.uniform(-0.1, 0.1, HID)
    p[W2:B2] = rng.uniform(-1, 1, HID * HID) * np.sqrt(6.0 / HID)
    p[B2:W3] = rng.uniform(-0.1, 0.1, HID)
    p[W3:B3] = rng.uniform(-0.2, 0.2, HID)
    p[B3] = 0.05
    return p


def forward(p: np.ndarray, X: np.ndarray, exact: bool = False):
    """o and the activations.  exact=False rounds inputs, weights and both
    activations to BF16 (as the kernels); exact=True is plain float64."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    q = (lambda a: np.asarray(a, dtype=np.float64)) if exact else (lambda a: bf16(a).astype(np.float64))
    # biases stay FP32 in the kernels
    fb = (lambda a: np.asarray(a, dtype=np.float64)) if exact else \
        (lambda a: np.asarray(a, dtype=np.float32).astype(np.float64))
    b1, b2, b3 = fb(b1), fb(b2), float(fb(b3))
    X = q(X)
    z1 = X @ q(w1) + b1
    h1 = np.maximum(z1, 0.0)
    h1 = h1 if exact else bf16(h1.astype(np.float32)).astype(np.float64)
    z2 = h1 @ q(w2) + b2
    h2 = np.maximum(z2, 0.0)
    h2 = h2 if exact else bf16(h2.astype(np.float32)).astype(np.float64)
    o = h2 @ q(w3) + b3
    return o, (X, h1, h2)


def backward(p: np.ndarray, X: np.ndarray, go: np.ndarray, exact: bool = False) -> np.ndarray:
    """d(sum go*o)/d(params) with the kernels' BF16 points (gz2 and gz1 are
    rounded because they feed MFMA) or exactly in float64."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    _, (Xq, h1, h2) = forward(p, X, exact)
    go = np.asarray(go, dtype=np.float64)
    if exact:
        w2q, w3q = w2.astype(np.float64), w3.astype(np.float64)
        gz2 = go[:, None] * w3q[None, :] * (h2 > 0)
        gz1 = (gz2 @ w2q.T) * (h1 > 0)
    else:
        w2q, w3q = bf16(w2).astype(np.float64), bf16(w3).astype(np.float64)
        gz2 = bf16((go[:, None] * w3q[None, :] * (h2 > 0)).astype(np.float32)).astype(np.float64)
        gz1 = bf16(((gz2 @ w2q.T) * (h1 > 0)).astype(np.float32)).astype(np.float64)
    g = np.zeros(NPARAMS, dtype=np.float64)
    g[W1:B1] = (Xq.T @ gz1).ravel()
    g[B1:W2] = gz1.sum(0)
    g[W2:B2] = (h1.T @ gz2).ravel()
    g[B2:W3] = gz2.sum(0)
    g[W3:B3] = h2.T @ go
    g[B3] = go.sum()
    return g


def jacobi_den(d: np.ndarray, theta: float) -> np.ndarray:
    """The reference's clamped Jacobi denominator (davidson.cpp:166-177)."""
    den = np.asarray(d, dtype=np.float64) - theta
    return np.where(np.abs(den) < 1e-10, np.where(den >= 0, 1e-10, -1e-10), den)


def features(diag: np.ndarray, occ: np.ndarray, conn: np.ndarray, theta: float, nsites: int) -> np.ndarray:
    """The 32 inputs per state: diag (Re H_ii), occ (n x 4 occupation
    fractions), conn (fraction of off-diagonal groups reaching a basis state
    with a non-zero element)."""
    d = np.asarray(diag, dtype=np.float64)
    den = jacobi_den(d, theta)
    f = np.zeros((d.size, IN), dtype=np.float32)
    f[:, 0] = np.tanh((d - theta) * 0.25)
    f[:, 1] = np.clip(1.0 / den, -4.0, 4.0) * 0.25
    f[:, 2] = theta / nsites
    f[:, 3] = d / nsites
    f[:, 4] = conn
    f[:, 5:9] = occ
    f[:, 9] = np.tanh(d - theta)
    return bf16(f)


def occupations(states: np.ndarray, nsites: int, two_bit: bool) -> np.ndarray:
    s = np.asarray(states, dtype=np.uint64)
    occ = np.zeros((s.size, 4), dtype=np.float64)
    for q in range(nsites):
        v = ((s >> np.uint64(2 * q)) & np.uint64(3)) if two_bit else ((s >> np.uint64(q)) & np.uint64(1))
        for k in range(4):
            occ[:, k] += (v == k)
    return (occ / nsites).astype(np.float32)


def pair_loss(diag, theta, o, r, ts, u):
    """L = 1 - |<ts, P_u t>|^2 / (|P_u t|^2 |ts|^2), t = (1/den + o) r, and
    dL/do.  ts must be orthogonal to u."""
    w = 1.0 / jacobi_den(diag, theta) + np.asarray(o, dtype=np.float64)
    t = w * r
    A = np.vdot(ts, t)
    c = np.vdot(u, t)
    S = np.sum(w * w * np.abs(r) ** 2)
    B = S - abs(c) ** 2
    C = np.vdot(ts, ts).real
    L = 1.0 - abs(A) ** 2 / (B * C)
    reA = (np.conj(A) * np.conj(ts) * r).real
    reC = (np.conj(c) * np.conj(u) * r).real
    g = -2.0 * reA / (B * C) + 2.0 * abs(A) ** 2 / (B * B * C) * (w * np.abs(r) ** 2 - reC)
    return L, g
