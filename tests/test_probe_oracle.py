"""Pins of the probe network's CPU oracle (oracle/probe_mlp.py): closed
forms, BF16 rounding on known bit patterns, and finite differences of the
float64 loss against the analytic gradient."""
import numpy as np

from oracle import probe_mlp as PM


def _params(D, H, seed, scale=0.5):
    rng = np.random.default_rng(seed)
    p = np.zeros(PM.param_count(D, H), np.float32)
    o = PM.offsets(D, H)
    p[slice(*o["W1"])] = rng.standard_normal(H * D) * scale / np.sqrt(D)
    p[slice(*o["b1"])] = rng.standard_normal(H) * 0.1
    p[slice(*o["W2"])] = rng.standard_normal(H * H) * scale / np.sqrt(H)
    p[slice(*o["b2"])] = rng.standard_normal(H) * 0.1
    p[slice(*o["w3"])] = rng.standard_normal(H) * scale / np.sqrt(H)
    p[slice(*o["b3"])] = 0.3
    p[slice(*o["mu"])] = rng.standard_normal(D) * 0.2
    p[slice(*o["isg"])] = 1.0 + rng.random(D)
    return p


def test_bf16_round_to_nearest_even():
    u = np.array([0x3F800000, 0x3F808000, 0x3F818000, 0x3F807FFF, 0x3F80C000, 0xBF808001, 0x7F7FFFFF],
                 np.uint32)
    want = np.array([0x3F800000, 0x3F800000, 0x3F820000, 0x3F800000, 0x3F810000, 0xBF810000, 0x7F800000],
                    np.uint32)
    got = PM.bf16(u.view(np.float32)).view(np.uint32)
    assert np.array_equal(got, want)


def test_gelu_closed_forms():
    assert PM.gelu(0.0) == 0.0
    assert abs(PM.gelu_grad(0.0) - 0.5) < 1e-15
    z = np.array([8.0, -8.0])
    assert np.allclose(PM.gelu(z), [8.0, 0.0], atol=1e-12)
    # Phi(1) = 0.8413447460685429 (normal CDF)
    assert abs(PM.gelu(1.0) - 0.8413447460685429) < 1e-15
    zz = np.linspace(-4, 4, 41)
    h = 1e-6
    num = (PM.gelu(zz + h) - PM.gelu(zz - h)) / (2 * h)
    assert np.allclose(num, PM.gelu_grad(zz), atol=1e-8)


def test_gradient_matches_finite_differences():
    D, H, n = 7, 6, 5
    p = _params(D, H, 1)
    rng = np.random.default_rng(2)
    raw = rng.standard_normal((n, D)).astype(np.float32)
    t = rng.standard_normal(n)
    g = PM.grads(p, raw, t, D, H, bf16_emul=False)
    k = PM.trained_count(D, H)
    assert g.shape == (k,)
    p64 = p.astype(np.float64)

    def L(q):
        # float64 evaluation of the same network (parameters not rounded to f32)
        gg = PM.unpack(q, D, H)
        X = (raw.astype(np.float64) - gg["mu"]) * gg["isg"]
        H1 = PM.gelu(X @ gg["W1"].T + gg["b1"])
        H2 = PM.gelu(H1 @ gg["W2"].T + gg["b2"])
        y = H2 @ gg["w3"] + gg["b3"][0]
        return np.mean((y - t) ** 2)

    eps = 1e-6
    for i in np.linspace(0, k - 1, 60).astype(int):
        a, b = p64.copy(), p64.copy()
        a[i] += eps
        b[i] -= eps
        num = (L(a) - L(b)) / (2 * eps)
        assert abs(num - g[i]) <= 1e-6 + 1e-5 * abs(num), (i, num, g[i])


def test_gradient_with_bf16_close_to_exact():
    # the BF16 roundings perturb the gradient by about their relative size
    D, H, n = 12, 16, 64
    p = _params(D, H, 3)
    rng = np.random.default_rng(4)
    raw = rng.standard_normal((n, D)).astype(np.float32)
    t = rng.standard_normal(n)
    g0 = PM.grads(p, raw, t, D, H, bf16_emul=False)
    g1 = PM.grads(p, raw, t, D, H, bf16_emul=True)
    assert np.linalg.norm(g1 - g0) <= 0.03 * np.linalg.norm(g0)


def test_adam_first_step_closed_form():
    # from zero moments the bias-corrected step is lr * g / (|g| + eps)
    g = np.array([0.5, -2.0, 1e-3, 0.0], np.float32)
    p0 = np.array([1.0, 1.0, 1.0, 1.0, 7.0], np.float32)
    m0 = np.zeros(5, np.float32)
    p1, m1, v1 = PM.adam(p0, m0, m0, g, 1, lr=0.01)
    want = p0[:4] - 0.01 * g / (np.abs(g) + 1e-8)
    assert np.allclose(p1[:4], want, rtol=1e-6, atol=1e-7)
    assert p1[4] == 7.0
    assert np.allclose(m1[:4], 0.1 * g) and np.allclose(v1[:4], 0.001 * g * g)
