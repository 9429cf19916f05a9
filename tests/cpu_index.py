"""CPU construction of small Puck indexes for tests.

Centers come from a seeded numpy Lloyd k-means (the reference's training is
seeded by time(0) and cannot be reproduced, so codebooks are this module's
own).  Cell assignment and filter-PQ encoding go through the oracle, which
restates HierarchicalClusterIndex::assign / PuckIndex::puck_assign.
"""
import numpy as np

from oracle import oracle as O


def kmeans_np(x, k, iters=8, seed=0):
    x = np.asarray(x, np.float32)
    rng = np.random.Generator(np.random.PCG64(seed))
    c = x[rng.choice(x.shape[0], size=k, replace=x.shape[0] < k)].copy()
    xn = (x * x).sum(1)
    for _ in range(iters):
        d = xn[:, None] - 2.0 * x @ c.T + (c * c).sum(1)[None, :]
        a = d.argmin(1)
        cnt = np.bincount(a, minlength=k)
        s = np.zeros_like(c)
        np.add.at(s, a, x)
        nz = cnt > 0
        c[nz] = s[nz] / cnt[nz, None]
        empty = np.where(~nz)[0]
        if len(empty):
            c[empty] = x[rng.choice(x.shape[0], size=len(empty))]
    return c.astype(np.float32)


def train_centers(x, C, F, iters=8, seed=0):
    """Coarse k-means, fine k-means on residuals, fine stored by descending norm
    (hierarchical_cluster_index.cpp:882-985, one outer iteration)."""
    coarse = kmeans_np(x, C, iters, seed)
    d = ((x[:, None, :] - coarse[None, :, :]) ** 2).sum(-1) if x.shape[0] * C < 4e7 else None
    if d is None:
        a = np.concatenate([(((x[s:s + 4096, None, :] - coarse[None]) ** 2).sum(-1)).argmin(1)
                            for s in range(0, x.shape[0], 4096)])
    else:
        a = d.argmin(1)
    res = x - coarse[a]
    fine = kmeans_np(res, F, iters, seed + 1)
    norms = np.array([O.dot(f, f) / np.float32(2) for f in fine], np.float32)
    order = np.argsort(-np.sqrt(norms), kind="stable")
    return coarse, np.ascontiguousarray(fine[order])


def train_codebook(residual, nsq, seed=0, iters=6):
    n, d = residual.shape
    lsq = d // nsq
    cb = np.zeros((nsq, 256, lsq), np.float32)
    for m in range(nsq):
        cb[m] = kmeans_np(residual[:, m * lsq:(m + 1) * lsq], 256, iters, seed + 17 * m)
    return cb


def build_index(x, C, F, nsq, whether_norm=False, seed=0, train_n=None):
    """Returns dict(coarse, fine, codebook, cell_assign, records, raw)."""
    x = np.asarray(x, np.float32)
    tn = x.shape[0] if train_n is None else min(train_n, x.shape[0])
    xt = x[:tn]
    coarse, fine = train_centers(xt, C, F, seed=seed)
    cell, _ = O.assign_cells(coarse, fine, x)
    c_id, f_id = cell // F, cell % F
    res = x - coarse[c_id] - fine[f_id]
    codebook = train_codebook(res[:tn], nsq, seed=seed + 3)
    records = O.encode_filter(coarse, fine, codebook, nsq, x, cell)
    return dict(coarse=coarse, fine=fine, codebook=codebook, cell_assign=cell, records=records,
                raw=x, whether_norm=whether_norm, nsq=nsq)


def build_index_pq(x, C, F, nsq, pq_nsq, whether_norm=False, seed=0, train_n=None):
    """A whether_pq = true index: build_index's filter PQ, then a second
    codebook of pq_nsq subspaces trained on the residuals left after the
    filter codewords (puck_index.cpp:836-871), and both quantizers' records
    (po_encode_puck).  Adds pq_codebook and pq_records."""
    ix = build_index(x, C, F, nsq, whether_norm, seed, train_n)
    x = ix["raw"]
    tn = x.shape[0] if train_n is None else min(train_n, x.shape[0])
    cell, coarse, fine, cb = ix["cell_assign"], ix["coarse"], ix["fine"], ix["codebook"]
    lsq = x.shape[1] // nsq
    res = x - coarse[cell // F] - fine[cell % F]
    codes = ix["records"][:, 4:].astype(np.int64)
    dec = np.concatenate([cb[m][codes[:, m]] for m in range(nsq)], axis=1).reshape(x.shape[0], nsq * lsq)
    res1 = (res - dec).astype(np.float32)
    pcb = train_codebook(res1[:tn], pq_nsq, seed=seed + 7)
    frec, prec = O.encode_puck(coarse, fine, cb, nsq, pcb, pq_nsq, x, cell)
    assert np.array_equal(frec, ix["records"])
    ix.update(pq_codebook=pcb, pq_records=prec, pq_nsq=pq_nsq)
    return ix


def build_index_with_centers(x, coarse, fine, nsq, whether_norm=False, seed=0):
    """Index over given coarse / fine centers (codebook trained on the residuals)."""
    x = np.asarray(x, np.float32)
    coarse = np.ascontiguousarray(coarse, np.float32)
    fine = np.ascontiguousarray(fine, np.float32)
    F = fine.shape[0]
    cell, _ = O.assign_cells(coarse, fine, x)
    res = x - coarse[cell // F] - fine[cell % F]
    codebook = train_codebook(res, nsq, seed=seed + 3)
    records = O.encode_filter(coarse, fine, codebook, nsq, x, cell)
    return dict(coarse=coarse, fine=fine, codebook=codebook, cell_assign=cell, records=records,
                raw=x, whether_norm=whether_norm, nsq=nsq)


def oracle_index(ix):
    oi = O.OracleIndex(ix["coarse"], ix["fine"], ix["codebook"], ix["cell_assign"],
                       ix["records"], ix["raw"], ix["whether_norm"])
    if "pq_records" in ix:
        oi.set_pq(ix["pq_codebook"], ix["pq_records"])
    return oi


def near_tie_encode_input(seed, nsq, lsq, n):
    """Filter-PQ encode input on which the order of the short dot products
    decides the codes: codewords and points a few ulps around one center, with
    zero coarse and fine centers (residual = point), so pn + cn - 2 dot cancels
    to rounding noise.  Returns (coarse, fine, codebook [nsq*256, lsq], x, cell)."""
    rng = np.random.default_rng(seed)
    d = nsq * lsq
    center = rng.uniform(1.0, 100.0, d).astype(np.float32)
    sp = np.spacing(center)
    cb = center.reshape(nsq, 1, lsq) + rng.integers(-8, 9, (nsq, 256, lsq)).astype(np.float32) * sp.reshape(nsq, 1, lsq)
    x = center + rng.integers(-64, 65, (n, d)).astype(np.float32) * sp
    z = np.zeros((1, d), np.float32)
    return (z, z.copy(), np.ascontiguousarray(cb.astype(np.float32).reshape(-1, lsq)), x.astype(np.float32),
            np.zeros(n, np.uint32))
