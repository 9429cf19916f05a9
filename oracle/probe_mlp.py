"""CPU restatement of the learned probe's network (test infrastructure only:
imported by tests/ and nothing on the product path).

The probe has no counterpart in the reference; this oracle states its
mathematics -- the forward pass, the mean-squared-error loss, its gradient
by the chain rule, and one Adam step -- in float64, with the BF16 rounding
of the GPU kernels (weights, layer inputs and the pre-activation gradients
that enter a matrix product) emulated at the same points when `bf16=True`.

Pinned by: finite differences of the float64 loss (tests/test_probe_oracle.py),
closed forms of GELU / GELU' and of an Adam step from zero moments, and
BF16 round-to-nearest-even on known bit patterns.

Packed parameter layout (D inputs, H hidden units):
    W1 [H x D], b1 [H], W2 [H x H], b2 [H], w3 [H], b3 [1], mu [D], 1/sigma [D]
"""
import numpy as np
from scipy.special import erf


def offsets(D, H):
    o = {}
    o["W1"] = (0, H * D)
    o["b1"] = (o["W1"][1], o["W1"][1] + H)
    o["W2"] = (o["b1"][1], o["b1"][1] + H * H)
    o["b2"] = (o["W2"][1], o["W2"][1] + H)
    o["w3"] = (o["b2"][1], o["b2"][1] + H)
    o["b3"] = (o["w3"][1], o["w3"][1] + 1)
    o["mu"] = (o["b3"][1], o["b3"][1] + D)
    o["isg"] = (o["mu"][1], o["mu"][1] + D)
    return o


def trained_count(D, H):
    return offsets(D, H)["b3"][1]


def param_count(D, H):
    return offsets(D, H)["isg"][1]


def unpack(p, D, H):
    o = offsets(D, H)
    g = {k: np.asarray(p[a:b]) for k, (a, b) in o.items()}
    g["W1"] = g["W1"].reshape(H, D)
    g["W2"] = g["W2"].reshape(H, H)
    return g


def bf16(x):
    """float32 -> bfloat16 -> float32, round to nearest even."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def _r(x, on):
    return bf16(np.asarray(x, np.float32)).astype(np.float64) if on else np.asarray(x, np.float64)


def forward(p, raw, D, H, bf16_emul=True):
    """Returns y [n] and the intermediates of the pass."""
    g = unpack(np.asarray(p, np.float32), D, H)
    xs = (np.asarray(raw, np.float32) - g["mu"].astype(np.float32)) * g["isg"].astype(np.float32)
    X = _r(xs, bf16_emul)
    W1, W2, w3 = _r(g["W1"], bf16_emul), _r(g["W2"], bf16_emul), _r(g["w3"], bf16_emul)
    Z1 = X @ W1.T + g["b1"].astype(np.float64)
    H1 = _r(gelu(Z1), bf16_emul)
    Z2 = H1 @ W2.T + g["b2"].astype(np.float64)
    H2 = _r(gelu(Z2), bf16_emul)
    y = H2 @ w3 + float(g["b3"][0])
    return y, dict(X=X, W1=W1, W2=W2, w3=w3, Z1=Z1, H1=H1, Z2=Z2, H2=H2)


def loss(p, raw, t, D, H, bf16_emul=True):
    y, _ = forward(p, raw, D, H, bf16_emul)
    return float(np.mean((y - np.asarray(t, np.float64)) ** 2))


def grads(p, raw, t, D, H, bf16_emul=True):
    """Gradient of the mean squared error w.r.t. the trained block (packed)."""
    y, c = forward(p, raw, D, H, bf16_emul)
    n = y.shape[0]
    dy = 2.0 * (y - np.asarray(t, np.float64)) / n
    gw3 = dy @ c["H2"]
    gb3 = np.array([dy.sum()])
    dZ2 = _r(dy[:, None] * c["w3"][None, :] * gelu_grad(c["Z2"]), bf16_emul)
    gW2 = dZ2.T @ c["H1"]
    gb2 = dZ2.sum(0)
    dH1 = dZ2 @ c["W2"]
    dZ1 = _r(dH1 * gelu_grad(c["Z1"]), bf16_emul)
    gW1 = dZ1.T @ c["X"]
    gb1 = dZ1.sum(0)
    return np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, gb3])


def adam(p, m, v, g, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step on the trained block (float32 state, as the kernel)."""
    p, m, v, g = (np.asarray(a, np.float32).copy() for a in (p, m, v, g))
    k = g.shape[0]
    m[:k] = np.float32(b1) * m[:k] + np.float32(1 - b1) * g
    v[:k] = np.float32(b2) * v[:k] + np.float32(1 - b2) * g * g
    c1 = np.float32(1 - b1 ** step)
    c2 = np.float32(1 - b2 ** step)
    p[:k] -= np.float32(lr) * (m[:k] / c1) / (np.sqrt(v[:k] / c2) + np.float32(eps))
    return p, m, v
