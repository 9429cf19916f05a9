"""GPU index construction: train coarse / fine centers and the filter-PQ
codebook, assign every point to its cell and encode its filter record.

Follows HierarchicalClusterIndex::train (hierarchical_cluster_index.cpp:882-985:
coarse k-means, fine k-means on the coarse residuals, fine centers stored by
descending norm) and PuckIndex::train (puck_index.cpp:764-874: per-subspace
256-center k-means on the cell residuals).  The Lloyd assignment step of every
k-means run is the engine's GEMM + argmin kernel (puckgpu_kmeans_assign, the
nearest_center formula of kmeans.cpp:221-258); the centroid means are
fixed-order torch reductions, so a build is reproducible run to run.  The
reference seeds its sampling and initialisation with time(0), so trained
codebooks are not reproducible against it; these are seeded.  Empty clusters
take half of the most populated ones (kmeans.cpp:112-147).
"""
import torch


def cluster_sums(x, a, k):
    """Per-cluster sums of the rows of x [n, d] (assignments a [n], long) and
    counts, in a fixed order: rows stably sorted by cluster, a float64 prefix
    sum, differences at the segment ends.  (index_add_ on a CUDA tensor adds
    with float atomics in no fixed order, so trained centers, and with them
    the index and its exact ties, would differ from run to run.)"""
    n, d = x.shape
    cnt = torch.bincount(a, minlength=k)
    order = torch.argsort(a, stable=True)
    cs = torch.cumsum(x[order].double(), 0)
    pad = torch.cat([torch.zeros((1, d), dtype=torch.float64, device=x.device), cs])
    ends = torch.cumsum(cnt, 0)
    sums = pad[ends] - pad[ends - cnt]
    return sums, cnt


def kmeans(ctx, x, k, niter=10, seed=0):
    """Lloyd k-means on a CUDA float32 tensor x [n, d]; returns [k, d]."""
    n, d = x.shape
    if k >= n:
        raise ValueError("better to have fewer clusters than points")
    g = torch.Generator(device=x.device)
    g.manual_seed(seed)
    c = x[torch.randperm(n, generator=g, device=x.device)[:k]].clone().contiguous()
    for _ in range(niter):
        a = ctx.kmeans_assign(c, x).long()
        sums, cnt = cluster_sums(x, a, k)
        nz = cnt > 0
        c[nz] = (sums[nz] / cnt[nz, None].double()).to(c.dtype)
        empty = torch.nonzero(~nz).flatten()
        if empty.numel():
            # split the most populated clusters (kmeans_reassign_empty, kmeans.cpp:112-147)
            donors = torch.argsort(cnt, descending=True)[: empty.numel()]
            eps = torch.randn((empty.numel(), d), generator=g, device=x.device) * 1e-7
            eps = eps * c[donors].norm(dim=1, keepdim=True)
            c[empty] = c[donors] - eps
            c[donors] = c[donors] + eps
        c = c.contiguous()
    return c


def train_centers(ctx, xs, C, F, niter=10, seed=0):
    """Coarse centers, fine (residual) centers stored by descending norm."""
    coarse = kmeans(ctx, xs, C, niter, seed)
    a = ctx.kmeans_assign(coarse, xs).long()
    res = (xs - coarse[a]).contiguous()
    fine = kmeans(ctx, res, F, niter, seed + 1)
    half = (fine * fine).sum(1) / 2
    order = torch.argsort(-torch.sqrt(half), stable=True)
    return coarse, fine[order].contiguous()


def train_codebook(ctx, coarse, fine, xs, nsq, niter=10, seed=7):
    """Filter-PQ codebook [nsq, 256, lsq] on the cell residuals of xs."""
    d = xs.shape[1]
    lsq = d // nsq
    F = fine.shape[0]
    cell = ctx.assign_cells(coarse, fine, xs).long()
    res = xs - coarse[cell // F] - fine[cell % F]
    cb = torch.empty((nsq, 256, lsq), dtype=torch.float32, device=xs.device)
    for m in range(nsq):
        cb[m] = kmeans(ctx, res[:, m * lsq:(m + 1) * lsq].contiguous(), 256, niter, seed + 31 * m)
    return cb


def train_pq_codebook(ctx, coarse, fine, codebook, nsq, xs, pq_nsq, niter=10, seed=13):
    """Second-stage codebook [pq_nsq, 256, d / pq_nsq] (whether_pq = true) on
    what the filter codewords leave of the cell residuals of xs: the
    quantizers are trained in turn, each on the residual its predecessors
    leave (PuckIndex::train, puck_index.cpp:836-871); the filter part is taken
    at the codes the encoder picks."""
    d = xs.shape[1]
    F = fine.shape[0]
    flsq, plsq = d // nsq, d // pq_nsq
    cell = ctx.assign_cells(coarse, fine, xs)
    rec = ctx.encode_filter(coarse, fine, codebook, nsq, xs, cell)
    cl = cell.long()
    res = (xs - coarse[cl // F]) - fine[cl % F]
    codes = rec[:, 4:].long()
    cbf = codebook.reshape(nsq, 256, flsq)
    dec = torch.stack([cbf[m][codes[:, m]] for m in range(nsq)], 1).reshape(xs.shape[0], d)
    res = (res - dec).contiguous()
    cb = torch.empty((pq_nsq, 256, plsq), dtype=torch.float32, device=xs.device)
    for m in range(pq_nsq):
        cb[m] = kmeans(ctx, res[:, m * plsq:(m + 1) * plsq].contiguous(), 256, niter, seed + 31 * m)
    return cb


def build(ctx, x, C, F, nsq, whether_norm, train_n=262144, niter=10, seed=0, pq_nsq=None):
    """Train, assign and encode x (CUDA [N, d]); returns the index arrays
    (device tensors in the reference file layouts) ready for GpuIndex.  With
    pq_nsq, also the second-stage PQ of a whether_pq = true index
    (pq_codebook, pq_records)."""
    N = x.shape[0]
    g = torch.Generator(device=x.device)
    g.manual_seed(seed + 99)
    sel = torch.randperm(N, generator=g, device=x.device)[: min(train_n, N)]
    xs = x[sel].contiguous()
    coarse, fine = train_centers(ctx, xs, C, F, niter, seed)
    codebook = train_codebook(ctx, coarse, fine, xs, nsq, niter, seed + 7)
    cb = codebook.reshape(-1).contiguous()
    cell = ctx.assign_cells(coarse, fine, x)
    rec = ctx.encode_filter(coarse, fine, cb, nsq, x, cell)
    out = dict(coarse=coarse, fine=fine, codebook=cb, cell_assign=cell, records=rec, raw=x,
               whether_norm=whether_norm, nsq=nsq)
    if pq_nsq:
        pcb = train_pq_codebook(ctx, coarse, fine, cb, nsq, xs, pq_nsq, niter, seed + 13).reshape(-1).contiguous()
        out.update(pq_codebook=pcb, pq_records=ctx.encode_pq(coarse, fine, cb, nsq, rec, pcb, pq_nsq, x, cell),
                   pq_nsq=pq_nsq)
    return out
