"""The 10M x 96 workload at its real size: the plain search, the batched
serving stream and the learned probe on one engine-built index, each checked
against the oracle.

Workload: 10,000,000 unit-norm vectors (4,096 normalised Gaussian means,
sigma 0.08, renormalised; seed 2), 10,000 queries from the same distribution
(seed 3); cosine evaluated as L2 on normalised vectors.  Index: 2,000 coarse x
2,000 fine centers, filter PQ nsq 48 (sub-vectors of 2 floats), trained,
assigned and encoded on the GPU by the engine.  filter_topk 1,000, top-k 10.

* Plain (puck_index.cpp:559-600): the whole batch at search_coarse_count 1
  (the operating point of recall@10 >= 0.95) and 4; a 500-query sample equals
  the oracle's result counts, ids and distance bits; the budget scan at a
  uniform budget on a 300-query sample equals the oracle's.
* Serving: 100,000 queries (seed 4) in batches of floor(exp(U[0, ln 4096])),
  per-query top-k in {1, 10, 100} and search_coarse_count in 20..400,
  filter_topk 1,100: the graph server returns, for every query, the bits of
  one eager batch call, and on a 300-query sample the oracle's results.
* Learned probe: trained on 100,000 extra queries (seed 5) with the engine's
  own labels and kernels; the learned search of the 10,000 queries equals, on
  a 300-query sample, the oracle's budget scan with the budgets the network
  gave; and at the budgets the network spends, its recall is above that of the
  same mean budget given to every query (the network's own gain).
"""
import sys
import time

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(1500)]


def _log(msg):
    # straight to the terminal: a long fixture shows its progress under capture
    print(f"[10M x 96] {msg}", file=sys.__stderr__, flush=True)


@pytest.fixture(scope="module")
def w10m():
    import torch
    from puck_amd import datasets, train
    from puck_amd.engine import Context, GpuIndex

    N, d, C, F, nsq, nq = 10_000_000, 96, 2000, 2000, 48, 10_000
    dev = torch.device("cuda:0")
    t0 = time.time()
    x = datasets.unit_norm_mixture(N, d, 4096, seed=2, means_seed=2000, backend="torch", device=dev)
    q = datasets.unit_norm_mixture(nq, d, 4096, seed=3, means_seed=2000, backend="torch", device=dev)
    ctx = Context(0)
    _log("building the index on the GPU")
    ix = train.build(ctx, x, C, F, nsq, whether_norm=True)
    torch.cuda.synchronize()
    _log(f"data + train + assign + encode {time.time() - t0:.1f}s")
    gi = GpuIndex(ctx, ix["coarse"], ix["fine"], ix["codebook"], ix["cell_assign"], ix["records"], ix["raw"], True)
    t0 = time.time()
    h = {k: ix[k].cpu().numpy() for k in ("coarse", "fine", "codebook", "cell_assign", "records")}
    oi = O.OracleIndex(h["coarse"], h["fine"], h["codebook"], h["cell_assign"].view(np.uint32), h["records"],
                       ix["raw"].cpu().numpy(), True)
    _log(f"oracle index {time.time() - t0:.1f}s")
    return dict(x=x, q=q, ctx=ctx, gi=gi, h=h, qh=q.cpu().numpy(), oi=oi, ix=ix, N=N, F=F, nq=nq)


def _same(Dg, Ig, Rg, Do, Io, Ro, rows, what):
    assert np.array_equal(Rg[rows], Ro), what
    for j, k in enumerate(rows):
        m = int(Ro[j])
        assert np.array_equal(Ig[k, :m], Io[j, :m]), f"{what}: query {k}"
        assert np.array_equal(Dg[k, :m].view(np.uint32), Do[j, :m].view(np.uint32)), f"{what}: query {k}"


def _host(D, I, R):
    return D.cpu().numpy(), I.cpu().numpy().view(np.uint32), R.cpu().numpy().view(np.uint32)


def test_plain_search_10m_x_96_matches_oracle(w10m):
    import torch
    from puck_amd.metrics import exact_topk, recall_with_ties

    ftk, topk = 1000, 10
    q, gi, qh, oi, nq, F = (w10m[k] for k in ("q", "gi", "qh", "oi", "nq", "F"))
    rng = np.random.default_rng(11)
    gt_i, gt_d = exact_topk(w10m["x"], q, topk)
    gt_i, gt_d = gt_i.cpu().numpy(), gt_d.cpu().numpy()
    for scc, ns in ((1, 500), (4, 200)):
        t0 = time.time()
        Dg, Ig, Rg = _host(*gi.search(q, topk, scc, ftk))
        rec, _ = recall_with_ties(gt_i, gt_d, Ig.astype(np.int64), topk)
        _log(f"plain scc={scc}: {nq} queries {time.time() - t0:.2f}s, recall@10 {rec:.4f}")
        qs = np.sort(rng.choice(nq, ns, replace=False))
        Do, Io, Ro, _ = oi.search_batch(qh[qs], topk, scc, ftk)
        _same(Dg, Ig, Rg, Do, Io, Ro, qs, f"plain scc={scc}")
        _log(f"plain scc={scc}: {ns}-query sample equals the oracle")
        if scc == 1:
            assert rec >= 0.95
    # the budget scan (the learned form's search) at a uniform budget
    nf = np.full(nq, 128, np.uint32)
    Dg, Ig, Rg = _host(*gi.search_batch(q, topk, 1, ftk, fine_budget=nf))
    qs = np.sort(rng.choice(nq, 300, replace=False))
    Do, Io, Ro = oi.search_batch_pq(qh[qs], topk, 1, ftk, fine_budget=nf[qs])
    _same(Dg, Ig, Rg, Do, Io, Ro, qs, "budget 128")
    _log("budget scan (128 cells): 300-query sample equals the oracle")


def test_serving_stream_10m_x_96(w10m):
    import torch
    from puck_amd import datasets
    from puck_amd.engine import Server

    gi, oi = w10m["gi"], w10m["oi"]
    nt, d = 100_000, 96
    dev = w10m["q"].device
    qs_all = datasets.unit_norm_mixture(nt, d, 4096, seed=4, means_seed=2000, backend="torch", device=dev)
    rng = np.random.default_rng(4)
    sizes, tot = [], 0
    while tot < nt:
        b = min(int(np.floor(np.exp(rng.uniform(0.0, np.log(4096))))), nt - tot)
        sizes.append(b)
        tot += b
    topk = rng.choice([1, 10, 100], nt).astype(np.uint32)
    scc = rng.integers(20, 401, nt).astype(np.uint32)
    tk = torch.as_tensor(topk.view(np.int32), device=dev)
    sc = torch.as_tensor(scc.view(np.int32), device=dev)
    D = torch.empty((nt, 100), dtype=torch.float32, device=dev)
    I = torch.empty((nt, 100), dtype=torch.int32, device=dev)
    R = torch.empty(nt, dtype=torch.int32, device=dev)
    t0 = time.time()
    sv = Server(gi, 100, 400, 1100, max_batch=4096)
    s0 = 0
    for b in sizes:
        sv.search(qs_all[s0:s0 + b], tk[s0:s0 + b], sc[s0:s0 + b], out=(D[s0:s0 + b], I[s0:s0 + b], R[s0:s0 + b]))
        s0 += b
    torch.cuda.synchronize()
    sv.free()
    _log(f"graph server: {nt} queries in {len(sizes)} batches {time.time() - t0:.1f}s (graph captures included)")
    Dg, Ig, Rg = _host(D, I, R)
    De, Ie, Re = _host(*gi.search_batch(qs_all, 100, 400, 1100, topk=tk, scc=sc))
    _same(Dg, Ig, Rg, De, Ie, Re, np.arange(nt), "server vs eager")
    rows = np.sort(np.random.default_rng(8).choice(nt, 300, replace=False))
    Do, Io, Ro = oi.search_batch_pq(qs_all[rows].cpu().numpy(), 100, 400, 1100, topk=topk[rows], scc_arr=scc[rows])
    _same(Dg, Ig, Rg, Do, Io, Ro, rows, "server vs oracle")
    _log("graph server equals the eager batch call on every query and the oracle on a 300-query sample")


@pytest.fixture(scope="module")
def probe10m(w10m):
    import torch
    from puck_amd import datasets
    from puck_amd import probe as PR
    from puck_amd.metrics import exact_topk

    t0 = time.time()
    qt = datasets.unit_norm_mixture(100_000, 96, 4096, seed=5, means_seed=2000, backend="torch",
                                    device=w10m["q"].device)
    gtt, _ = exact_topk(w10m["x"], qt, 10)
    cell = w10m["ix"]["cell_assign"].view(torch.int32)
    probe, info = PR.fit(w10m["ctx"], w10m["gi"], qt, gtt, cell, 1, 1000, 10, epochs=6, log=None)
    torch.cuda.synchronize()
    _log(f"probe trained on 100,000 queries in {time.time() - t0:.1f}s: {info}")
    return probe


def test_learned_probe_10m_x_96(w10m, probe10m):
    import torch
    from puck_amd.engine import search_learned
    from puck_amd.metrics import exact_topk, recall_with_ties

    q, gi, qh, oi, nq, F = (w10m[k] for k in ("q", "gi", "qh", "oi", "nq", "F"))
    gt_i, gt_d = exact_topk(w10m["x"], q, 10)
    gt_i, gt_d = gt_i.cpu().numpy(), gt_d.cpu().numpy()
    rows = np.sort(np.random.default_rng(12).choice(nq, 300, replace=False))
    learned = []
    for margin in (0.0, 0.5):
        Dg, Ig, Rg, Bg = search_learned(gi, probe10m, q, 10, 1, 1000, margin=margin, fallback_frac=0.5)
        Dg, Ig, Rg = _host(Dg, Ig, Rg)
        B = Bg.cpu().numpy().view(np.uint32)
        Do, Io, Ro = oi.search_batch_pq(qh[rows], 10, 1, 1000, fine_budget=B[rows])
        _same(Dg, Ig, Rg, Do, Io, Ro, rows, f"learned margin {margin}")
        rec, _ = recall_with_ties(gt_i, gt_d, Ig.astype(np.int64), 10)
        mb = float(np.where(B >= F, F, B).mean())
        learned.append((margin, mb, rec))
        _log(f"learned margin {margin}: recall@10 {rec:.4f}, mean budget {mb:.1f}, "
             f"{(B >= F).mean():.4f} of queries fall back to the walk; 300-query sample equals the oracle")
    # the network's own gain: the same mean budget given to every query
    for margin, mb, rec in learned:
        b = int(round(mb))
        nf = torch.full((nq,), b, dtype=torch.int32, device=q.device)
        _, Iu, _ = gi.search_batch(q, 10, 1, 1000, fine_budget=nf)
        ru, _ = recall_with_ties(gt_i, gt_d, Iu.cpu().numpy().astype(np.int64), 10)
        _log(f"margin {margin}: learned recall {rec:.4f} vs uniform budget {b}: {ru:.4f}")
        assert rec > ru, f"margin {margin}: the network does not beat its mean budget given uniformly"
