"""Measurement helpers: exact brute-force top-k on the GPU and recall with
distance ties (ann-benchmarks benchmark/plotting/metrics.py:15-40)."""
import numpy as np
import torch


@torch.no_grad()
def exact_topk(x, q, k, extra=8, qchunk=4096, bchunk=1 << 21):
    """Exact squared-L2 top-(k+extra) of q [nq, d] in x [N, d] (CUDA tensors):
    FP32 GEMM candidates per base block, then an FP64 recheck that decides the
    order.  Returns (ids int64 [nq, k+extra], dists float64 [nq, k+extra])."""
    kk = k + extra
    nq, N = q.shape[0], x.shape[0]
    # x may be uint8 (integer features): blocks are converted to float32
    xn = (x.double() ** 2).sum(1).float() if N <= (1 << 20) else torch.cat(
        [(x[s:s + bchunk].double() ** 2).sum(1).float() for s in range(0, N, bchunk)])
    out_i = torch.empty((nq, kk), dtype=torch.int64, device=q.device)
    out_d = torch.empty((nq, kk), dtype=torch.float64, device=q.device)
    for qs in range(0, nq, qchunk):
        qq = q[qs:qs + qchunk]
        cand_i, cand_d = [], []
        for bs in range(0, N, bchunk):
            xb = x[bs:bs + bchunk].float()
            d = torch.addmm(xn[bs:bs + bchunk][None, :], qq, xb.T, beta=1.0, alpha=-2.0)
            v, i = torch.topk(d, min(kk, xb.shape[0]), dim=1, largest=False)
            cand_i.append(i + bs)
            cand_d.append(v)
        ci = torch.cat(cand_i, 1)
        v, j = torch.topk(torch.cat(cand_d, 1), min(2 * kk, ci.shape[1]), dim=1, largest=False)
        ci = torch.gather(ci, 1, j)
        exact = ((x[ci].double() - qq.double()[:, None, :]) ** 2).sum(-1)
        order = torch.argsort(exact, dim=1, stable=True)[:, :kk]
        out_i[qs:qs + qchunk] = torch.gather(ci, 1, order)
        out_d[qs:qs + qchunk] = torch.gather(exact, 1, order)
    return out_i, out_d


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
