"""GPU index construction: train coarse / fine centers and the PQ codebooks,
assign every point to its cell and encode its records.

Follows HierarchicalClusterIndex::train (hierarchical_cluster_index.cpp:
882-985) through puckgpu_train_centers -- rounds of coarse k-means on the
points less their best fine centers, fine k-means on the points less their
new coarse centers, keep-best with the Lloyd limit growing by 10, fine
centers stored by descending norm -- and PuckIndex::train (puck_index.cpp:
764-874) through puckgpu_kmeans: every quantizer in turn, every subspace a
k-means++ run of 256 centers on what the earlier subspaces and quantizers
leave of the cell residuals.  Every k-means is the engine's (nearest_center
assignments, the reference's centroid arithmetic, empty-cluster splits).  The
reference seeds each run with time(0); these runs are seeded, so a build is
reproducible.
"""
import torch


def train_centers(ctx, xs, C, F, rounds=10, niter=30, seed=0, mean_mode=0):
    """Coarse centers [C, d] and fine centers [F, d] (descending norm) of the
    CUDA sample xs."""
    coarse, fine, _, _, _ = ctx.train_centers(xs, C, F, iterations=rounds, niter=niter, seed=seed,
                                              mean_mode=mean_mode)
    return coarse, fine


def train_quantizers(ctx, coarse, fine, xs, nsqs, niter=30, seed=7, mean_mode=0):
    """PuckIndex::train's codebooks (puck_index.cpp:807-871): the residuals
    (x - c) - f of the sample's cells; for each quantizer (nsqs: the filter
    PQ's subspace count, then the second-stage PQ's) and each subspace in
    turn, 256 centers by k-means++ and then the subspace residual less the
    center each point was assigned.  Returns one [nsq, 256, d / nsq]
    codebook per quantizer."""
    d = xs.shape[1]
    F = fine.shape[0]
    cell = ctx.assign_cells(coarse, fine, xs).long()
    res = (xs - coarse[cell // F]) - fine[cell % F]
    books = []
    for qi, nsq in enumerate(nsqs):
        lsq = d // nsq
        cb = torch.empty((nsq, 256, lsq), dtype=torch.float32, device=xs.device)
        for m in range(nsq):
            sub = res[:, m * lsq:(m + 1) * lsq].contiguous()
            c, a, _, _, _ = ctx.kmeans(sub, 256, niter=niter, init_pp=True, seed=seed + 1000 * qi + m,
                                       mean_mode=mean_mode)
            cb[m] = c
            res[:, m * lsq:(m + 1) * lsq] = sub - c[a.long()]
        books.append(cb)
    return books


def train_codebook(ctx, coarse, fine, xs, nsq, niter=30, seed=7, mean_mode=0):
    """Filter-PQ codebook [nsq, 256, d / nsq] (whether_pq = false)."""
    return train_quantizers(ctx, coarse, fine, xs, [nsq], niter, seed, mean_mode)[0]


def build(ctx, x, C, F, nsq, whether_norm, train_n=262144, niter=30, seed=0, pq_nsq=None, rounds=10,
          mean_mode=0):
    """Train, assign and encode x (CUDA [N, d]); returns the index arrays
    (device tensors in the reference file layouts) ready for GpuIndex.  With
    pq_nsq, also the second-stage PQ of a whether_pq = true index
    (pq_codebook, pq_records).  rounds: kmeans_iterations_count; niter: the
    first Lloyd limit of every k-means."""
    N = x.shape[0]
    g = torch.Generator(device=x.device)
    g.manual_seed(seed + 99)
    sel = torch.randperm(N, generator=g, device=x.device)[: min(train_n, N)]
    xs = x[sel].contiguous()
    coarse, fine = train_centers(ctx, xs, C, F, rounds, niter, seed, mean_mode)
    books = train_quantizers(ctx, coarse, fine, xs, [nsq] + ([pq_nsq] if pq_nsq else []), niter, seed + 7,
                             mean_mode)
    cb = books[0].reshape(-1).contiguous()
    cell = ctx.assign_cells(coarse, fine, x)
    rec = ctx.encode_filter(coarse, fine, cb, nsq, x, cell)
    out = dict(coarse=coarse, fine=fine, codebook=cb, cell_assign=cell, records=rec, raw=x,
               whether_norm=whether_norm, nsq=nsq)
    if pq_nsq:
        pcb = books[1].reshape(-1).contiguous()
        out.update(pq_codebook=pcb, pq_records=ctx.encode_pq(coarse, fine, cb, nsq, rec, pcb, pq_nsq, x, cell),
                   pq_nsq=pq_nsq)
    return out


def build_chunked(ctx, x, C, F, nsq, whether_norm, train_n=10_000_000, pq_train_n=1_000_000, niter=30, seed=0,
                  rounds=10, mean_mode=0, chunk=1 << 22, log=None):
    """build() for point sets whose float rows do not fit at once: x is a
    CUDA [N, d] float32 or uint8 tensor (uint8: integer features, which the
    index keeps as uint8 rows).  The centers are trained on train_n sampled
    points (train_points_count), the codebook on the first pq_train_n of them
    (pq_train_points_count, puck_index.cpp:40), then the points are assigned
    and encoded chunk by chunk.  Returns build()'s dict; raw is x itself.
    log(msg), when given, receives a line per step (and the library's
    training progress)."""
    log = log or (lambda m: None)
    ctx.set_progress(log)
    N = x.shape[0]
    g = torch.Generator(device=x.device)
    g.manual_seed(seed + 99)
    sel = torch.randperm(N, generator=g, device=x.device)[: min(train_n, N)]
    xs = x[sel].float().contiguous()
    del sel
    coarse, fine = train_centers(ctx, xs, C, F, rounds, niter, seed, mean_mode)
    log(f"centers trained on {xs.shape[0]} points")
    xq = xs[: min(pq_train_n, xs.shape[0])].contiguous()
    del xs
    torch.cuda.empty_cache()
    cb = train_quantizers(ctx, coarse, fine, xq, [nsq], niter, seed + 7, mean_mode)[0].reshape(-1).contiguous()
    log(f"filter codebook trained on {xq.shape[0]} points")
    del xq
    cell = torch.empty(N, dtype=torch.int32, device=x.device)
    rec = torch.empty((N, 4 + nsq), dtype=torch.uint8, device=x.device)
    for s0 in range(0, N, chunk):
        xc = x[s0:s0 + chunk].float().contiguous()
        cell[s0:s0 + xc.shape[0]] = ctx.assign_cells(coarse, fine, xc)
        rec[s0:s0 + xc.shape[0]] = ctx.encode_filter(coarse, fine, cb, nsq, xc, cell[s0:s0 + xc.shape[0]])
        del xc
        log(f"assigned and encoded {min(N, s0 + chunk)} of {N}")
    ctx.set_progress(None)
    return dict(coarse=coarse, fine=fine, codebook=cb, cell_assign=cell, records=rec, raw=x,
                whether_norm=whether_norm, nsq=nsq)
