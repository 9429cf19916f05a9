"""Training pipeline of the learned probe (labels, initialisation, Adam
loop) driving the device kernels through the C ABI.

Fine budget b of a query: in each coarse list the budget scan scores every
point of the b cells of smallest cell distance (include/puck_gpu.h,
puckgpu_query_params).  Label of a query: run the plain search (the
reference's walk) at the operating search_coarse_count; of the exact top-k
neighbours it returns, take the largest fine-budget rank of a neighbour's cell
(puckgpu_budget_ranks); the label is that rank (1 when it returns none).  A
budget >= the label scores those neighbours in full, and a neighbour among
the 10 nearest points stays in the top-10 once it is in the filter set.  It
leaves the filter set only if filter_topk scored points have smaller filter
distances: the plain walk prunes whole cells by the distance of their centers
(a heuristic for the points inside), so the scan can score points the walk
skipped.  On the test indexes every neighbour is found again; the labels rely
on that, it is not a guarantee.  The network regresses log(label).
"""
import numpy as np
import torch

from .engine import Probe


def init_params(D, H, mu, sigma, seed=0):
    """He-normal weights, zero biases, feature scaling (mu, 1/sigma)."""
    rng = np.random.default_rng(seed)
    parts = [rng.standard_normal(H * D) * np.sqrt(2.0 / D), np.zeros(H),
             rng.standard_normal(H * H) * np.sqrt(2.0 / H), np.zeros(H),
             rng.standard_normal(H) * np.sqrt(1.0 / H), np.zeros(1),
             np.asarray(mu, np.float64), 1.0 / np.maximum(np.asarray(sigma, np.float64), 1e-6)]
    return np.concatenate(parts).astype(np.float32)


@torch.no_grad()
def budget_labels(index, queries, gt_ids, cell_assign, scc, filter_topk, topk=10, chunk=16384):
    """Labels (int64 tensor [n]) as described in the module docstring.
    queries, gt_ids [n, >=topk] and cell_assign [N] (local-order cell ids,
    coarse * F + fine) are CUDA tensors."""
    n = queries.shape[0]
    lab = torch.empty(n, dtype=torch.int64, device=queries.device)
    ar = torch.arange(topk, device=queries.device)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        qc = queries[s:e].contiguous()
        _, I, R = index.search(qc, topk, scc, filter_topk)
        I = I.long()
        valid = ar[None, :] < R.long()[:, None]
        hit = (I[:, :, None] == gt_ids[s:e, None, :topk].long()).any(-1) & valid
        cells = cell_assign.long()[I.clamp(min=0, max=cell_assign.shape[0] - 1)]
        cells = torch.where(hit, cells, torch.full_like(cells, 0xffffffff)).to(torch.int64)
        cells32 = (cells & 0xffffffff).to(torch.int64)
        cells32 = torch.where(cells32 >= 2 ** 31, cells32 - 2 ** 32, cells32).to(torch.int32).contiguous()
        r = index.budget_ranks(qc, cells32, scc).long() & 0xffffffff
        r = torch.where(r == 0xffffffff, torch.zeros_like(r), r)
        lab[s:e] = r.max(1).values.clamp(min=1)
    return lab


def train(probe, feats, target, epochs=8, batch=1024, lr=1e-3, lr_final=1e-4, seed=0, log=None):
    """Adam over (feats [n, D], target [n]) CUDA tensors with a cosine
    learning-rate decay; returns the per-epoch mean training loss."""
    n = feats.shape[0]
    g = torch.Generator(device=feats.device)
    g.manual_seed(seed)
    steps = epochs * ((n + batch - 1) // batch)
    hist, k = [], 0
    for ep in range(epochs):
        perm = torch.randperm(n, generator=g, device=feats.device)
        tot, cnt = 0.0, 0
        for s in range(0, n, batch):
            idx = perm[s:s + batch]
            if idx.numel() < 2:
                continue
            xb = feats.index_select(0, idx).contiguous()
            yb = target.index_select(0, idx).contiguous()
            a = lr_final + 0.5 * (lr - lr_final) * (1 + np.cos(np.pi * k / max(1, steps - 1)))
            want = (s // batch) % 16 == 0
            loss = probe.train_step(xb, yb, lr=float(a), want_loss=want)
            if want:
                tot += loss
                cnt += 1
            k += 1
        hist.append(tot / max(cnt, 1))
        if log:
            log(f"probe epoch {ep}: mse {hist[-1]:.4f}")
    return hist


def fit(ctx, index, train_queries, gt_ids, cell_assign, scc, filter_topk, topk=10, hidden=256, nc=64, nf=64,
        epochs=8, batch=1024, lr=1e-3, seed=0, log=None):
    """Build a Probe for `index` from training queries (CUDA tensors): features
    and labels from the engine, then Adam.  Returns (probe, info)."""
    probe = Probe(ctx, index.dim, nc, nf, hidden)
    feats = probe.features(index, train_queries)
    lab = budget_labels(index, train_queries, gt_ids, cell_assign, scc, filter_topk, topk)
    target = torch.log(lab.double()).float().contiguous()
    mu = feats.double().mean(0).cpu().numpy()
    sd = feats.double().std(0).cpu().numpy()
    probe.set_params(init_params(probe.D, hidden, mu, sd, seed))
    hist = train(probe, feats, target, epochs, batch, lr, seed=seed, log=log)
    info = dict(n_train=int(train_queries.shape[0]), label_mean=float(lab.double().mean()),
                label_p50=float(lab.double().median()), label_p99=float(torch.quantile(lab.double(), 0.99)),
                mse_per_epoch=[round(h, 4) for h in hist], target_var=float(target.double().var()))
    return probe, info


def save(probe, path):
    """Writes the probe's shape and packed float32 parameters (W1, b1, W2, b2,
    w3, b3, feature mean, 1/std; the layout of puckgpu_probe_set_params) to
    an .npz file."""
    np.savez(path, shape=np.array([probe.dim, probe.nc, probe.nf, probe.hidden], np.int64),
             params=probe.get_params())


def load(ctx, path):
    """A Probe on ctx's device from a file written by save()."""
    with np.load(path, allow_pickle=False) as z:
        dim, nc, nf, hidden = (int(v) for v in z["shape"])
        params = np.ascontiguousarray(z["params"], np.float32)
    probe = Probe(ctx, dim, nc, nf, hidden)
    probe.set_params(params)
    return probe
