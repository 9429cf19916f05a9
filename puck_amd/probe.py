"""Training pipeline of the learned probe (labels, initialisation, Adam
loop) driving the device kernels through the C ABI.

Label of a query: the fine budget its fixed-budget search needs.  Run the
plain search (every fine cluster, the reference's walk) at the operating
search_coarse_count; of the exact top-k neighbours it returns, take the
largest rank of a neighbour's fine cluster in the query's ascending
fine-distance order; the label is that rank + 1 (1 when it returns none).
A budget >= the label walks the same cells up to that point, so the
neighbours the plain search finds are found again (the filter heap of a
truncated walk is a prefix state of the full walk's).  The network
regresses log(label).
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
def budget_labels(index, queries, fine_order, gt_ids, cell_assign, scc, filter_topk, topk=10, chunk=16384):
    """Labels (int64 tensor [n]) as described in the module docstring.
    queries, fine_order [n, F], gt_ids [n, >=topk] and cell_assign [N] are
    CUDA tensors; cell_assign holds local-order cell ids (coarse * F + fine)."""
    n, F = queries.shape[0], index.F
    lab = torch.empty(n, dtype=torch.int64, device=queries.device)
    ar = torch.arange(F, dtype=torch.int64, device=queries.device)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        _, I, R = index.search(queries[s:e].contiguous(), topk, scc, filter_topk)
        I = I.long()
        valid = ar[None, :topk] < R.long()[:, None]
        gt = gt_ids[s:e, :topk].long()
        hit = (I[:, :, None] == gt[:, None, :]).any(-1) & valid
        fo = fine_order[s:e].long()
        rank = torch.empty_like(fo)
        rank.scatter_(1, fo, ar[None, :].expand(e - s, F))
        fine_of = (cell_assign.long()[I.clamp(min=0, max=cell_assign.shape[0] - 1)] % F)
        r = torch.gather(rank, 1, fine_of)
        r = torch.where(hit, r + 1, torch.zeros_like(r))
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
    feats, forder = probe.features(index, train_queries, fine_order=True)
    lab = budget_labels(index, train_queries, forder, gt_ids, cell_assign, scc, filter_topk, topk)
    del forder
    target = torch.log(lab.double()).float().contiguous()
    mu = feats.double().mean(0).cpu().numpy()
    sd = feats.double().std(0).cpu().numpy()
    probe.set_params(init_params(probe.D, hidden, mu, sd, seed))
    hist = train(probe, feats, target, epochs, batch, lr, seed=seed, log=log)
    info = dict(n_train=int(train_queries.shape[0]), label_mean=float(lab.double().mean()),
                label_p50=float(lab.double().median()), label_p99=float(torch.quantile(lab.double(), 0.99)),
                mse_per_epoch=[round(h, 4) for h in hist], target_var=float(target.double().var()))
    return probe, info
