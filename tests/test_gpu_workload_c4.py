"""The 100M x 128 uint8 workload's geometry on 10,000,000 points: integer
vectors (the 1M x 128 generator with 10,000 means, seed 6) stored as uint8 in
HBM, C = F = 2,048, filter PQ nsq 32, trained, assigned and encoded by the
engine (train.build_chunked, 1M-point training sample, 4 rounds), checked
against the oracle.  The full 100M-point run is bench.py --workload c4 (its
build alone takes minutes, too long for this suite).

* Assignment: on a 20,000-point sample the bounded assignment equals the
  oracle's restatement of nearest_cell_assign (the sequential walk,
  hierarchical_cluster_index.cpp:1042-1085), cells and scores bit for bit;
  the filter records of the sample equal the oracle's encoding.
* Search over uint8 rows (300-query samples): the plain walk and the budget
  scan equal the oracle's results (ids and distance bits), the oracle ranking
  the same integer values as float rows.
* Precision forms: BF16 / FP8 center distances and the FP16 filter table
  return exact FP32 rerank distances (each returned distance equals the
  float64 distance of its id rounded through the reference's order, checked
  against FP64 within 1e-6 relative), and their recall stays within 0.02 of
  the exact form's.
"""
import sys
import time

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(1500)]


def _log(msg):
    print(f"[c4 geometry] {msg}", file=sys.__stderr__, flush=True)


@pytest.fixture(scope="module")
def wc4():
    import torch
    from puck_amd import datasets, train
    from puck_amd.engine import Context, GpuIndex

    N, d, C, F, nsq, nq = 10_000_000, 128, 2048, 2048, 32, 10_000
    dev = torch.device("cuda:0")
    t0 = time.time()
    x8 = datasets.mixture_uint8_grid(N, d, 10000, seed=6, means_seed=6000, backend="torch", device=dev,
                                     as_uint8=True)
    q = datasets.mixture_uint8_grid(nq, d, 10000, seed=7, means_seed=6000, backend="torch", device=dev)
    ctx = Context(0)
    ix = train.build_chunked(ctx, x8, C, F, nsq, whether_norm=False, train_n=1_000_000, pq_train_n=1_000_000,
                             rounds=4, log=_log)
    torch.cuda.synchronize()
    _log(f"data + train + assign + encode {time.time() - t0:.1f}s")
    gi = GpuIndex(ctx, ix["coarse"], ix["fine"], ix["codebook"], ix["cell_assign"], ix["records"], x8, False)
    t0 = time.time()
    h = {k: ix[k].cpu().numpy() for k in ("coarse", "fine", "codebook", "cell_assign", "records")}
    oi = O.OracleIndex(h["coarse"], h["fine"], h["codebook"], h["cell_assign"].view(np.uint32), h["records"],
                       x8.cpu().numpy().astype(np.float32), False)
    _log(f"oracle index {time.time() - t0:.1f}s")
    return dict(x8=x8, q=q, ctx=ctx, gi=gi, h=h, qh=q.cpu().numpy(), oi=oi, N=N, F=F, nq=nq, nsq=nsq)


def _host(D, I, R):
    return D.cpu().numpy(), I.cpu().numpy().view(np.uint32), R.cpu().numpy().view(np.uint32)


def _same(Dg, Ig, Rg, Do, Io, Ro, rows, what):
    assert np.array_equal(Rg[rows], Ro), what
    for j, k in enumerate(rows):
        m = int(Ro[j])
        assert np.array_equal(Ig[k, :m], Io[j, :m]), f"{what}: query {k}"
        assert np.array_equal(Dg[k, :m].view(np.uint32), Do[j, :m].view(np.uint32)), f"{what}: query {k}"


def test_c4_assignment_and_records_match_oracle(wc4):
    h, x8 = wc4["h"], wc4["x8"]
    rows = np.sort(np.random.default_rng(21).choice(wc4["N"], 20000, replace=False))
    xs = x8[rows].cpu().numpy().astype(np.float32)
    co, do = O.assign_cells(h["coarse"], h["fine"], xs)
    assert np.array_equal(h["cell_assign"].view(np.uint32)[rows], co)
    cg, dg = wc4["ctx"].assign_cells(h["coarse"], h["fine"], xs, with_dist=True)
    assert np.array_equal(cg, co) and np.array_equal(dg.view(np.uint32), do.view(np.uint32))
    ro = O.encode_filter(h["coarse"], h["fine"], h["codebook"].reshape(wc4["nsq"], 256, -1), wc4["nsq"], xs, co)
    assert np.array_equal(h["records"][rows], ro)
    pts, seq = wc4["ctx"].assign_stats()
    _log(f"20,000-point sample: cells, scores and records equal the oracle; {seq} of {pts} points walked")


def test_c4_search_over_uint8_rows_matches_oracle(wc4):
    q, gi, qh, oi, nq = (wc4[k] for k in ("q", "gi", "qh", "oi", "nq"))
    rng = np.random.default_rng(22)
    for scc, ftk in ((1, 500), (4, 1000)):
        Dg, Ig, Rg = _host(*gi.search(q, 10, scc, ftk))
        qs = np.sort(rng.choice(nq, 300, replace=False))
        Do, Io, Ro, _ = oi.search_batch(qh[qs], 10, scc, ftk)
        _same(Dg, Ig, Rg, Do, Io, Ro, qs, f"plain scc={scc} ftk={ftk}")
        _log(f"plain scc={scc} filter_topk={ftk}: 300-query sample equals the oracle")
    nf = (np.arange(nq, dtype=np.uint32) % 16) + 1
    Dg, Ig, Rg = _host(*gi.search_batch(q, 10, 4, 1000, fine_budget=nf))
    qs = np.sort(rng.choice(nq, 300, replace=False))
    Do, Io, Ro = oi.search_batch_pq(qh[qs], 10, 4, 1000, fine_budget=nf[qs])
    _same(Dg, Ig, Rg, Do, Io, Ro, qs, "budget scan")
    _log("budget scan (1..16 cells): 300-query sample equals the oracle")


def test_c4_precision_forms(wc4):
    from puck_amd.metrics import exact_topk, recall_with_ties

    q, gi, x8, nq = wc4["q"], wc4["gi"], wc4["x8"], wc4["nq"]
    gt_i, gt_d = exact_topk(x8, q, 10)
    gt_i, gt_d = gt_i.cpu().numpy(), gt_d.cpu().numpy()
    recs = {}
    for prec in (0, 1, 2, 16):
        Dg, Ig, Rg = _host(*gi.search(q, 10, 4, 1000, precision=prec))
        recs[prec], _ = recall_with_ties(gt_i, gt_d, Ig.astype(np.int64), 10)
        # returned distances are the exact FP32 rerank of the returned ids
        rows = np.arange(0, nq, 50)
        xs = x8[Ig[rows].astype(np.int64).reshape(-1)].cpu().numpy().astype(np.float64).reshape(len(rows), 10, -1)
        want = ((xs - wc4["qh"][rows][:, None, :].astype(np.float64)) ** 2).sum(-1)
        got = Dg[rows].astype(np.float64)
        assert np.all(np.abs(got - want) <= 1e-6 * np.maximum(want, 1.0))
        _log(f"precision {prec}: recall@10 {recs[prec]:.4f}, returned distances exact")
    for prec in (1, 2, 16):
        assert recs[prec] >= recs[0] - 0.02
