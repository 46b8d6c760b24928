"""CPU restatement of the learned Davidson preconditioner (test infrastructure).

Only tests/ may import this module; the library never calls it.  The learned
preconditioner has no counterpart in the reference, whose Davidson corrects
with the Jacobi diagonal (src/solvers/davidson.cpp:166-177): this file
restates the library's own definition in numpy, with the BF16 rounding at
the same points as the GPU kernels, so that the kernels' forward pass,
parameter gradients, features, correction terms and pair loss can be checked.

Pinned by: float32->bfloat16 rounding against torch's conversion; the exact
(float64, unrounded) backward against central finite differences of the
exact forward; the pair-loss gradient against finite differences and the
loss at its two ends (0 for t along t*, 1 for K t orthogonal to K t*); the
correction terms against a dense solve (z - w and z - w + w2 are the first-
and second-order Neumann expansions of (H - theta)^-1 r, so their errors
fall as the second and third power of the off-diagonal scale); the network
with a zero output layer giving z - w/2.

Network: x (32) -> relu(x W1 + b1) (128) -> relu(h1 W2 + b2) (128) -> h2 W3 + b3 (2).
Parameter block (float32): W1[32][128], b1[128], W2[128][128], b2[128],
W3[128][2], b3[2], row index = input.

Correction: z = J r, w = J (H z - diag z), w2 = J (H w - diag w), J =
1/(H_ii - theta) clamped as the reference clamps it; t = z + s1 w + s2 w2,
s1 = -1/2 + tanh(a)/2, s2 = tanh(b)/2, (a, b) = network output per state
(a zero output gives the damped first-order form z - w/2).
"""
from __future__ import annotations

import numpy as np

IN, HID, OUT = 32, 128, 2
W1, B1 = 0, IN * HID
W2 = B1 + HID
B2 = W2 + HID * HID
W3 = B2 + HID
B3 = W3 + HID * OUT
NPARAMS = B3 + OUT


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
            p[W3:B3].reshape(HID, OUT), p[B3:NPARAMS])


def init_params(seed: int) -> np.ndarray:
    """Any parameter block works with the checks; this one is not the library's
    initialisation (which uses std::mt19937 on the host)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAMS, dtype=np.float32)
    p[W1:B1] = rng.uniform(-1, 1, IN * HID) * np.sqrt(6.0 / IN)
    p[B1:W2] = rng.uniform(-0.1, 0.1, HID)
    p[W2:B2] = rng.uniform(-1, 1, HID * HID) * np.sqrt(6.0 / HID)
    p[B2:W3] = rng.uniform(-0.1, 0.1, HID)
    p[W3:B3] = rng.uniform(-0.2, 0.2, HID * OUT)
    p[B3:NPARAMS] = (0.05, -0.03)
    return p


def forward(p: np.ndarray, X: np.ndarray, exact: bool = False):
    """o (n x 2) and the activations.  exact=False rounds inputs, weights and
    both activations to BF16 (as the kernels); exact=True is plain float64."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    q = (lambda a: np.asarray(a, dtype=np.float64)) if exact else (lambda a: bf16(a).astype(np.float64))
    # biases stay FP32 in the kernels
    fb = (lambda a: np.asarray(a, dtype=np.float64)) if exact else \
        (lambda a: np.asarray(a, dtype=np.float32).astype(np.float64))
    b1, b2, b3 = fb(b1), fb(b2), fb(b3)
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
    """d(sum go*o)/d(params), go (n x 2), with the kernels' BF16 points (gz2
    and gz1 are rounded because they feed MFMA) or exactly in float64."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    _, (Xq, h1, h2) = forward(p, X, exact)
    go = np.asarray(go, dtype=np.float64).reshape(-1, OUT)
    if exact:
        w2q, w3q = w2.astype(np.float64), w3.astype(np.float64)
        gz2 = (go @ w3q.T) * (h2 > 0)
        gz1 = (gz2 @ w2q.T) * (h1 > 0)
    else:
        w2q, w3q = bf16(w2).astype(np.float64), bf16(w3).astype(np.float64)
        gz2 = bf16(((go @ w3q.T) * (h2 > 0)).astype(np.float32)).astype(np.float64)
        gz1 = bf16(((gz2 @ w2q.T) * (h1 > 0)).astype(np.float32)).astype(np.float64)
    g = np.zeros(NPARAMS, dtype=np.float64)
    g[W1:B1] = (Xq.T @ gz1).ravel()
    g[B1:W2] = gz1.sum(0)
    g[W2:B2] = (h1.T @ gz2).ravel()
    g[B2:W3] = gz2.sum(0)
    g[W3:B3] = (h2.T @ go).ravel()
    g[B3:NPARAMS] = go.sum(0)
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


def neumann_terms(diag, theta, r, apply_h):
    """z = J r, w = J (H z - diag z) and w2 = J (H w - diag w); apply_h(x) = H x."""
    d = np.asarray(diag, dtype=np.float64)
    J = 1.0 / jacobi_den(diag, theta)
    z = J * r
    w = J * (apply_h(z) - d * z)
    w2 = J * (apply_h(w) - d * w)
    return z, w, w2


def coefficients(o):
    """(s1, s2) per state from the network output o (n x 2)."""
    o = np.asarray(o, dtype=np.float64).reshape(-1, OUT)
    return -0.5 + 0.5 * np.tanh(o[:, 0]), 0.5 * np.tanh(o[:, 1])


def combine(o, z, w, w2):
    s1, s2 = coefficients(o)
    return z + s1 * w + s2 * w2


def pair_loss(o, z, w, w2, b, u, theta, apply_h):
    """Training loss of one pair and dL/do (n x 2).  With P = 1 - u u^H (|u| = 1),
    K = P (H - theta) P and t = P (z + s1 w + s2 w2):
        L = 1 - |<b, K t>|^2 / (|K t|^2 |b|^2),
    b = K t* for the accurate correction t*.  apply_h(x) = H x."""
    o = np.asarray(o, dtype=np.float64).reshape(-1, OUT)
    ta, tb = np.tanh(o[:, 0]), np.tanh(o[:, 1])
    P = lambda x: x - u * np.vdot(u, x)
    K = lambda x: P(apply_h(x) - theta * x)
    t = P(combine(o, z, w, w2))
    q = K(t)
    a = np.vdot(b, q)
    Q = np.vdot(q, q).real
    Bn = np.vdot(b, b).real
    L = 1.0 - abs(a) ** 2 / (Q * Bn)
    g = -2.0 * a * b / (Q * Bn) + 2.0 * abs(a) ** 2 * q / (Q * Q * Bn)
    h = K(P(g))
    grad = np.stack([(np.conj(h) * w).real * 0.5 * (1.0 - ta * ta),
                     (np.conj(h) * w2).real * 0.5 * (1.0 - tb * tb)], axis=1)
    return L, grad
