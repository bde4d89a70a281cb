"""CPU restatement (float64 numpy) of the troubled-interface classifier:
stencil features, MLP forward pass, weighted binary cross-entropy and its
gradient by back-propagation, and the forward pass with BF16 rounding of
the inputs, weights and hidden activations (the solver's matrix-core
inference).  TEST INFRASTRUCTURE ONLY: used by tests/ as the checker of the
GPU kernels (include/gks_learn.h); the solver never calls it.

The classifier has no counterpart in the reference; the formulas restate
the parameter layout and features documented in include/gks_learn.h and
high-order-solver_amd/csrc/gks2d_learn_device.hpp.
"""
import numpy as np

NIN, NH = 16, 32
NPARAM = NIN * NH + NH + NH * NH + NH + NH + 1


def unpack(P):
    P = np.asarray(P, dtype=np.float64)
    W1 = P[0:512].reshape(NIN, NH)
    b1 = P[512:544]
    W2 = P[544:1568].reshape(NH, NH)
    b2 = P[1568:1600]
    W3 = P[1600:1632]
    b3 = P[1632]
    return W1, b1, W2, b2, W3, b3


def forward(P, X):
    W1, b1, W2, b2, W3, b3 = unpack(P)
    X = np.asarray(X, dtype=np.float64)
    a1 = X @ W1 + b1
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ W2 + b2
    h2 = np.maximum(a2, 0.0)
    return h2 @ W3 + b3, (a1, h1, a2, h2)


def loss_grad(P, X, y, w=None):
    """Mean over the batch of w * BCE(logit, y), and its gradient."""
    X = np.asarray(X, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    w = np.ones_like(y) if w is None else np.asarray(w, dtype=np.float64)
    n = len(y)
    z, (a1, h1, a2, h2) = forward(P, X)
    W1, b1, W2, b2, W3, b3 = unpack(P)
    loss = np.sum(w * (np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z))))) / n
    gz = w * (1.0 / (1.0 + np.exp(-z)) - y) / n
    g2 = np.outer(gz, W3) * (a2 > 0)
    g1 = (g2 @ W2.T) * (a1 > 0)
    grad = np.concatenate([(X.T @ g1).ravel(), g1.sum(0), (h1.T @ g2).ravel(), g2.sum(0), h2.T @ gz,
                           [gz.sum()]])
    return loss, grad


def bf16(a):
    """Round float32 values to the nearest BF16 (ties to even), as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def forward_bf16(P, X):
    """Forward pass as the matrix cores run it: BF16 inputs, weights and
    hidden activations, float accumulation, float biases."""
    W1, b1, W2, b2, W3, b3 = unpack(P)
    x = bf16(X).astype(np.float64)
    h1 = bf16(np.maximum(x @ bf16(W1).astype(np.float64) + b1.astype(np.float32), 0).astype(np.float32))
    h2 = bf16(np.maximum(h1.astype(np.float64) @ bf16(W2).astype(np.float64) + b2.astype(np.float32), 0)
              .astype(np.float32))
    return h2.astype(np.float64) @ bf16(W3).astype(np.float64) + np.float32(b3)


def features(q):
    """16 features of stencils q [n, 2, 5] ((rho, p) x 5 cells), computed in
    float32 with the device's operation order."""
    q = np.asarray(q, dtype=np.float64)
    n = q.shape[0]
    out = np.zeros((n, 16), dtype=np.float32)
    f32 = np.float32
    for v in range(2):
        s = q[:, v, :].astype(np.float32)
        m = ((((s[:, 0] + s[:, 1]) + s[:, 2]) + s[:, 3]) + s[:, 4]) * f32(0.2)
        spread = s.max(axis=1) - s.min(axis=1)
        inv = np.where(spread > 0, f32(1.0) / np.where(spread > 0, spread, f32(1.0)), f32(0.0)).astype(np.float32)
        out[:, 5 * v:5 * v + 5] = (s - m[:, None]) * inv[:, None]
        out[:, 10 + v] = np.log10(spread / (np.abs(m) + f32(1e-30)) + f32(1e-16)) * f32(0.0625)
    return out
