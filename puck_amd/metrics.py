"""Measurement helpers: exact brute-force top-k on the GPU (the engine's own
exact search) and recall with distance ties (ann-benchmarks
benchmark/plotting/metrics.py:15-40)."""
import numpy as np
import torch


_CTX = {}


def _context(device):
    from .engine import Context
    dv = torch.device(device).index or 0
    if dv not in _CTX:
        _CTX[dv] = Context(dv)
    return _CTX[dv]


@torch.no_grad()
def exact_topk(x, q, k, extra=8, ctx=None):
    """Exact squared-L2 top-(k+extra) of q [nq, d] in x [N, d] (CUDA tensors,
    x float32 or uint8) by the engine's own exact search
    (puckgpu_exact_search: exact FP32 GEMM candidates per 65,536-point chunk,
    merged, ordered in FP64 by (distance, id)).  Returns (ids int64
    [nq, k+extra], dists float64 [nq, k+extra])."""
    ctx = ctx or _context(q.device)
    ids, dist = ctx.exact_search(x, q.contiguous(), k + extra, extra=8)
    return ids.long(), dist


def recall_with_ties(true_ids, true_dists, run_ids, count):
    """Mean recall@count; ties within 1e-6 of the count-th true distance count
    as hits (metrics.py:15-40).  Arrays are numpy."""
    nq = len(run_ids)
    rec = np.zeros(nq)
    for i in range(nq):
        td = true_dists[i]
        end = len(td)
        if end > count:
            anchor = td[count - 1]
            end = count
            for j in range(count, len(td)):
                if abs(anchor - td[j]) < 1e-6:
                    end = j + 1
                else:
                    break
        rec[i] = len(set(true_ids[i][:end].tolist()) & set(np.asarray(run_ids[i]).tolist()))
    return float(rec.mean() / count), rec
