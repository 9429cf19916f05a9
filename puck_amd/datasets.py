"""Seeded synthetic workloads for the Puck batched-search path.

Two generators, each available on the CPU (numpy, PCG64) and on the GPU
(torch CUDA generator, for the 10M-point sets):

* ``mixture_uint8_grid`` -- a mixture of isotropic Gaussians with means drawn
  uniformly in [0, mean_high]^d, sigma 25, clipped to [0, 255] and rounded:
  integer-valued float vectors (1M x 128, L2).
* ``unit_norm_mixture`` -- Gaussian means normalised to the unit sphere,
  points = mean + N(0, sigma^2 I), renormalised (10M x 96, cosine as L2).

Seeds are explicit arguments; the two backends draw different streams, so a
workload is identified by (generator, backend, seeds, sizes).
"""
import numpy as np


def mixture_uint8_grid(n, d=128, n_means=1000, seed=0, means_seed=1000, mean_high=160.0,
                       sigma=25.0, backend="numpy", device="cuda", chunk=1 << 20, as_uint8=False):
    """Rows are integers 0..255: float32, or uint8 with as_uint8 (the same
    values, a quarter of the bytes; the torch backend then never holds the
    float rows of more than one chunk)."""
    if backend == "numpy":
        rm = np.random.Generator(np.random.PCG64(means_seed))
        means = rm.uniform(0.0, mean_high, size=(n_means, d)).astype(np.float32)
        r = np.random.Generator(np.random.PCG64(seed))
        comp = r.integers(0, n_means, size=n)
        x = means[comp] + r.standard_normal((n, d), dtype=np.float32) * np.float32(sigma)
        return np.rint(np.clip(x, 0.0, 255.0)).astype(np.uint8 if as_uint8 else np.float32)
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(means_seed)
    means = torch.rand((n_means, d), generator=g, device=device) * mean_high
    g.manual_seed(seed)
    out = torch.empty((n, d), dtype=torch.uint8 if as_uint8 else torch.float32, device=device)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        comp = torch.randint(0, n_means, (e - s,), generator=g, device=device)
        x = means[comp] + torch.randn((e - s, d), generator=g, device=device) * sigma
        out[s:e] = torch.round(torch.clamp(x, 0.0, 255.0)).to(out.dtype)
    return out


def unit_norm_mixture(n, d=96, n_means=4096, seed=2, means_seed=2000, sigma=0.08,
                      backend="numpy", device="cuda", chunk=1 << 20):
    if backend == "numpy":
        rm = np.random.Generator(np.random.PCG64(means_seed))
        means = rm.standard_normal((n_means, d)).astype(np.float32)
        means /= np.linalg.norm(means, axis=1, keepdims=True)
        r = np.random.Generator(np.random.PCG64(seed))
        comp = r.integers(0, n_means, size=n)
        x = means[comp] + r.standard_normal((n, d), dtype=np.float32) * np.float32(sigma)
        x /= np.linalg.norm(x, axis=1, keepdims=True)
        return x.astype(np.float32)
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(means_seed)
    means = torch.randn((n_means, d), generator=g, device=device)
    means = means / means.norm(dim=1, keepdim=True)
    g.manual_seed(seed)
    out = torch.empty((n, d), dtype=torch.float32, device=device)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        comp = torch.randint(0, n_means, (e - s,), generator=g, device=device)
        x = means[comp] + torch.randn((e - s, d), generator=g, device=device) * sigma
        out[s:e] = x / x.norm(dim=1, keepdim=True)
    return out
