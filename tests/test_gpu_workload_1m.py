"""The 1M x 128 plain-Puck workload at its real size, built and searched by the
engine and checked against the oracle.

Workload: 1,000,000 integer-valued vectors in [0, 255] from a mixture of 1,000
isotropic Gaussians (means uniform in [0, 160]^128, sigma 25, clipped and
rounded); 10,000 queries from the same mixture; L2.  Index: 1,024 coarse x
1,024 fine centers from k-means (10 iterations on a 200k sample), filter PQ
nsq 32 with 256 codewords, all trained, assigned and encoded on the GPU.
Search: search_coarse_count 100, filter_topk 500, top-k 10, one batch of
10,000 queries (the batch size that takes the grouped query order).

Checked: cell assignments and filter records of a 2,000-point sample equal
the oracle's (po_assign_cells / po_encode_filter, bit for bit); a 500-query
sample of the batch returns the oracle's PuckIndex::search result count, ids
and distance bits (puck_index.cpp:559-600).  Recall@10 against the exact
brute force is printed.  The batched-serving parameters run on the same
index: the graph server over batches of floor(exp(U[0, ln 4096])) queries with
per-query top-k in {1, 10, 100} and search_coarse_count in 20..400
(filter_topk 1100) returns, for every query, the bits of one eager batch call
and, on a sample, the oracle's per-query results.
"""
import time

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _log(msg):
    print(f"[1M x 128] {msg}", flush=True)


@pytest.fixture(scope="module")
def w1m():
    import torch
    from puck_amd import datasets, train
    from puck_amd.engine import Context, GpuIndex

    N, d, C, F, nsq, nq = 1_000_000, 128, 1024, 1024, 32, 10_000
    dev = torch.device("cuda:0")
    t0 = time.time()
    x = datasets.mixture_uint8_grid(N, d, 1000, seed=0, means_seed=1000, backend="torch", device=dev)
    q = datasets.mixture_uint8_grid(nq, d, 1000, seed=1, means_seed=1000, backend="torch", device=dev)
    ctx = Context(0)
    ix = train.build(ctx, x, C, F, nsq, whether_norm=False, train_n=200_000, niter=10, rounds=3)
    torch.cuda.synchronize()
    _log(f"data + train + assign + encode {time.time() - t0:.1f}s")
    gi = GpuIndex(ctx, ix["coarse"], ix["fine"], ix["codebook"], ix["cell_assign"], ix["records"], ix["raw"], False)
    h = {k: ix[k].cpu().numpy() for k in ("coarse", "fine", "codebook", "cell_assign", "records")}
    xh = x.cpu().numpy()
    oi = O.OracleIndex(h["coarse"], h["fine"], h["codebook"], h["cell_assign"].view(np.uint32), h["records"], xh,
                       False)
    return dict(x=x, q=q, ctx=ctx, gi=gi, h=h, xh=xh, qh=q.cpu().numpy(), oi=oi, N=N, nsq=nsq, nq=nq)


def test_plain_search_1m_x_128_matches_oracle(w1m):
    import torch
    from puck_amd.metrics import exact_topk, recall_with_ties

    scc, ftk, topk = 100, 500, 10
    x, q, gi, h, xh, qh, oi = (w1m[k] for k in ("x", "q", "gi", "h", "xh", "qh", "oi"))
    N, nsq, nq = w1m["N"], w1m["nsq"], w1m["nq"]
    t0 = time.time()
    Dg, Ig, Rg = gi.search(q, topk, scc, ftk)
    torch.cuda.synchronize()
    _log(f"search of {nq} queries {time.time() - t0:.2f}s")
    Dg, Ig, Rg = Dg.cpu().numpy(), Ig.cpu().numpy().view(np.uint32), Rg.cpu().numpy().view(np.uint32)
    rng = np.random.default_rng(7)

    t0 = time.time()
    ps = np.sort(rng.choice(N, 2000, replace=False))
    cell_o, _ = O.assign_cells(h["coarse"], h["fine"], xh[ps])
    assert np.array_equal(h["cell_assign"].view(np.uint32)[ps], cell_o)
    rec_o = O.encode_filter(h["coarse"], h["fine"], h["codebook"], nsq, xh[ps], cell_o)
    assert np.array_equal(h["records"][ps], rec_o)
    _log(f"assignment and encoding sample equal the oracle ({time.time() - t0:.1f}s)")

    t0 = time.time()
    qs = np.sort(rng.choice(nq, 500, replace=False))
    Do, Io, Ro, _ = oi.search_batch(qh[qs], topk, scc, ftk)
    assert np.array_equal(Rg[qs], Ro)
    for j, k in enumerate(qs):
        m = int(Ro[j])
        assert np.array_equal(Ig[k, :m], Io[j, :m]), f"query {k}"
        assert np.array_equal(Dg[k, :m].view(np.uint32), Do[j, :m].view(np.uint32)), f"query {k}"
    _log(f"500-query sample equals the oracle ({time.time() - t0:.1f}s)")

    gt_i, gt_d = exact_topk(x, q, topk)
    rec, _ = recall_with_ties(gt_i.cpu().numpy(), gt_d.cpu().numpy(), Ig.astype(np.int64), topk)
    _log(f"recall@10 {rec:.4f}")
    assert rec > 0.5


def test_serving_parameters_1m_x_128(w1m):
    import torch
    from puck_amd.engine import Server

    q, gi, qh, oi, nq = (w1m[k] for k in ("q", "gi", "qh", "oi", "nq"))
    dev = q.device
    rng = np.random.default_rng(4)
    sizes, tot = [], 0
    while tot < nq:
        b = min(int(np.floor(np.exp(rng.uniform(0.0, np.log(4096))))), nq - tot)
        sizes.append(b)
        tot += b
    topk = rng.choice([1, 10, 100], nq).astype(np.uint32)
    scc = rng.integers(20, 401, nq).astype(np.uint32)
    tk = torch.as_tensor(topk.view(np.int32), device=dev)
    sc = torch.as_tensor(scc.view(np.int32), device=dev)
    D = torch.empty((nq, 100), dtype=torch.float32, device=dev)
    I = torch.empty((nq, 100), dtype=torch.int32, device=dev)
    R = torch.empty(nq, dtype=torch.int32, device=dev)
    t0 = time.time()
    sv = Server(gi, 100, 400, 1100, max_batch=4096)
    s0 = 0
    for b in sizes:
        sv.search(q[s0:s0 + b], tk[s0:s0 + b], sc[s0:s0 + b], out=(D[s0:s0 + b], I[s0:s0 + b], R[s0:s0 + b]))
        s0 += b
    torch.cuda.synchronize()
    sv.free()
    _log(f"graph server: {nq} queries in {len(sizes)} batches {time.time() - t0:.1f}s (graph captures included)")
    De, Ie, Re = gi.search_batch(q, 100, 400, 1100, topk=tk, scc=sc)
    Dg, Ig, Rg = D.cpu().numpy(), I.cpu().numpy().view(np.uint32), R.cpu().numpy().view(np.uint32)
    De, Ie, Re = De.cpu().numpy(), Ie.cpu().numpy().view(np.uint32), Re.cpu().numpy().view(np.uint32)
    assert np.array_equal(Rg, Re)
    for k in range(nq):
        m = int(Re[k])
        assert np.array_equal(Ig[k, :m], Ie[k, :m]), f"query {k}"
        assert np.array_equal(Dg[k, :m].view(np.uint32), De[k, :m].view(np.uint32)), f"query {k}"
    qs = np.sort(np.random.default_rng(8).choice(nq, 300, replace=False))
    Do, Io, Ro = oi.search_batch_pq(qh[qs], 100, 400, 1100, topk=topk[qs], scc_arr=scc[qs])
    assert np.array_equal(Rg[qs], Ro)
    for j, k in enumerate(qs):
        m = int(Ro[j])
        assert np.array_equal(Ig[k, :m], Io[j, :m]), f"query {k}"
        assert np.array_equal(Dg[k, :m].view(np.uint32), Do[j, :m].view(np.uint32)), f"query {k}"
    _log("graph server equals the eager batch call on every query and the oracle on a 300-query sample")
