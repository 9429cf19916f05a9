"""GPU tests of the learned probe through the C ABI: the network's forward
pass, gradients and Adam step against the CPU oracle (oracle/probe_mlp.py),
its features against the oracle's center distances, the fused learned
search against explicit budgets and the oracle, and the label property."""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import probe_mlp as PM
from puck_amd import datasets as DS
import cpu_index as CI

pytestmark = pytest.mark.gpu

D_, H_ = 224, 256   # 96 + 64 + 64 inputs, 256 hidden units


@pytest.fixture(scope="module")
def ctx():
    from puck_amd.engine import Context
    return Context(0)


@pytest.fixture(scope="module")
def idx_cos():
    x = DS.unit_norm_mixture(15000, 96, 300, seed=20, means_seed=21)
    q = DS.unit_norm_mixture(64, 96, 300, seed=22, means_seed=21)
    ix = CI.build_index(x, 40, 40, 48, whether_norm=True, train_n=6000)
    return ix, q


def _gpu_index(ctx, ix):
    from puck_amd.engine import GpuIndex
    return GpuIndex(ctx, ix["coarse"], ix["fine"], ix["codebook"].reshape(-1), ix["cell_assign"],
                    ix["records"], ix["raw"], ix["whether_norm"])


def _rand_params(D, H, seed):
    rng = np.random.default_rng(seed)
    mu = rng.standard_normal(D) * 0.1
    sd = 0.5 + rng.random(D)
    from puck_amd.probe import init_params
    p = init_params(D, H, mu, sd, seed)
    o = PM.offsets(D, H)
    p[slice(*o["b1"])] = rng.standard_normal(H) * 0.1
    p[slice(*o["b2"])] = rng.standard_normal(H) * 0.1
    p[slice(*o["b3"])] = 0.7
    return p


def _probe(ctx, p, nc=64, nf=64, dim=96, hidden=H_):
    from puck_amd.engine import Probe
    pr = Probe(ctx, dim, nc, nf, hidden)
    pr.set_params(p)
    return pr


def test_probe_forward_matches_oracle(ctx):
    p = _rand_params(D_, H_, 1)
    pr = _probe(ctx, p)
    rng = np.random.default_rng(2)
    for n in (1, 63, 300):
        raw = rng.standard_normal((n, D_)).astype(np.float32)
        y = pr.predict(raw)
        yo, _ = PM.forward(p, raw, D_, H_, bf16_emul=True)
        # FP32 accumulation vs float64: BF16 roundings of a few activations may differ by one unit
        assert np.max(np.abs(y - yo)) <= 2e-2 * max(1.0, np.max(np.abs(yo))), n
        assert np.mean(np.abs(y - yo)) <= 2e-3 * max(1.0, np.mean(np.abs(yo))), n
    assert np.array_equal(pr.get_params(), p)


def test_probe_gradients_match_oracle(ctx):
    p = _rand_params(D_, H_, 3)
    pr = _probe(ctx, p)
    rng = np.random.default_rng(4)
    n = 1000
    raw = rng.standard_normal((n, D_)).astype(np.float32)
    t = rng.standard_normal(n).astype(np.float32)
    loss, g = pr.train_step(raw, t, lr=None, want_grads=True)
    go = PM.grads(p, raw, t, D_, H_, bf16_emul=True)
    assert abs(loss - PM.loss(p, raw, t, D_, H_)) <= 1e-3 * max(1.0, loss)
    o = PM.offsets(D_, H_)
    for k in ("W1", "b1", "W2", "b2", "w3", "b3"):
        a, b = o[k]
        err = np.linalg.norm(g[a:b] - go[a:b])
        assert err <= 1e-2 * np.linalg.norm(go[a:b]) + 1e-6, (k, err, np.linalg.norm(go[a:b]))
    # no update without Adam parameters
    assert np.array_equal(pr.get_params(), p)


def test_probe_adam_step_matches_oracle(ctx):
    p = _rand_params(D_, H_, 5)
    pr = _probe(ctx, p)
    rng = np.random.default_rng(6)
    raw = rng.standard_normal((256, D_)).astype(np.float32)
    t = rng.standard_normal(256).astype(np.float32)
    _, g = pr.train_step(raw, t, lr=None, want_grads=True)
    pr.train_step(raw, t, lr=1e-3)
    got = pr.get_params()
    k = PM.trained_count(D_, H_)
    z = np.zeros(p.shape[0], np.float32)
    want, _, _ = PM.adam(p, z, z, g, 1, lr=1e-3)
    assert np.allclose(got[:k], want[:k], rtol=1e-6, atol=1e-7)
    assert np.array_equal(got[k:], p[k:])


def test_probe_training_reduces_loss(ctx):
    from puck_amd.probe import init_params
    rng = np.random.default_rng(7)
    n = 4096
    raw = rng.standard_normal((n, D_)).astype(np.float32)
    w = rng.standard_normal(D_) / np.sqrt(D_)
    t = (np.tanh(raw @ w) * 2.0 + 0.5 * (raw[:, 0] ** 2)).astype(np.float32)
    pr = _probe(ctx, init_params(D_, H_, raw.mean(0), raw.std(0), 0))
    first = pr.train_step(raw[:512], t[:512], lr=None)
    for ep in range(30):
        for s in range(0, n, 512):
            pr.train_step(raw[s:s + 512], t[s:s + 512], lr=2e-3, want_loss=False)
    last = pr.train_step(raw[:512], t[:512], lr=None)
    assert last < 0.1 * first, (first, last)


def _clamp_empty(row):
    # an empty cell's distance (the stationary sentinel sqrt(FLT_MAX)) becomes
    # the row's largest populated-cell distance (rows ascending), 0 if none
    row = np.array(row, np.float32)
    lim = np.float32(np.sqrt(np.finfo(np.float32).max, dtype=np.float32) / np.float32(2))
    for t in range(row.shape[0]):
        if row[t] >= lim:
            row[t] = row[t - 1] if t > 0 else np.float32(0)
    return row


def test_probe_features_match_oracle(ctx, idx_cos):
    ix, q = idx_cos
    gi = _gpu_index(ctx, ix)
    nc, nf = 16, 12
    pr = _probe(ctx, np.zeros(PM.param_count(96 + nc + nf, 32), np.float32), nc=nc, nf=nf, hidden=32)
    f, fo = pr.features(gi, q, fine_order=True)
    o = CI.oracle_index(ix).export()
    for k in range(q.shape[0]):
        qn = O.normalize(q[k])
        assert np.array_equal(f[k, :96].view(np.uint32), qn.view(np.uint32))
        cd = np.array([o["coarse_norms"][c] - O.dot(ix["coarse"][c], qn) for c in range(40)], np.float32)
        fd = np.array([-O.dot(ix["fine"][j], qn) for j in range(40)], np.float32)
        assert np.array_equal(f[k, 96:96 + nc].view(np.uint32), np.sort(cd)[:nc].view(np.uint32))
        c0 = int(np.argmin(cd))
        cell = (cd[c0] + o["stationary"].reshape(40, 40)[c0]) + fd      # float32, reference order
        assert np.array_equal(f[k, 96 + nc:].view(np.uint32), _clamp_empty(np.sort(cell)[:nf]).view(np.uint32))
        assert np.array_equal(fd[fo[k]], np.sort(fd))
        assert np.array_equal(np.sort(fo[k]), np.arange(40))


@pytest.fixture(scope="module")
def idx_mid():
    # integer data (exact distance ties) with 300 coarse and 120 fine clusters:
    # rows of 64 keys or more go through the candidate bound of the row top-k
    x = DS.mixture_uint8_grid(20000, 32, 300, seed=70, means_seed=71)
    q = DS.mixture_uint8_grid(48, 32, 300, seed=72, means_seed=71)
    ix = CI.build_index(x, 300, 120, 16, whether_norm=False, train_n=8000)
    return ix, q


def test_probe_features_rows_over_64(ctx, idx_mid):
    ix, q = idx_mid
    gi = _gpu_index(ctx, ix)
    C, F, d = 300, 120, 32
    for nc, nf in [(64, 64), (1, 1), (200, 100)]:
        pr = _probe(ctx, np.zeros(PM.param_count(d + nc + nf, 16), np.float32), nc=nc, nf=nf, dim=d, hidden=16)
        f = pr.features(gi, q)
        o = CI.oracle_index(ix).export()
        for k in range(q.shape[0]):
            cd = np.array([o["coarse_norms"][c] - O.dot(ix["coarse"][c], q[k]) for c in range(C)], np.float32)
            fd = np.array([-O.dot(ix["fine"][j], q[k]) for j in range(F)], np.float32)
            assert np.array_equal(f[k, d:d + nc].view(np.uint32), np.sort(cd)[:nc].view(np.uint32)), (nc, k)
            c0 = int(np.argmin(cd))
            cell = (cd[c0] + o["stationary"].reshape(C, F)[c0]) + fd
            want = _clamp_empty(np.sort(cell)[:nf])
            assert np.array_equal(f[k, d + nc:].view(np.uint32), want.view(np.uint32)), (nf, k)
            assert np.all(np.isfinite(f[k])) and np.abs(f[k]).max() < 1e18, k


@pytest.mark.parametrize("scc,ftk,topk", [(1, 100, 10), (16, 300, 10), (64, 500, 20)])
def test_search_rows_over_64_bitexact(ctx, idx_mid, scc, ftk, topk):
    ix, q = idx_mid
    gi = _gpu_index(ctx, ix)
    oi = CI.oracle_index(ix)
    Dg, Ig, Rg = gi.search(q, topk, scc, ftk)
    for k in range(q.shape[0]):
        o = oi.search_one(q[k], topk, scc, ftk)
        assert Rg[k] == o["n"], k
        assert np.array_equal(Ig[k, :o["n"]], o["ids"][:o["n"]]), k
        assert np.array_equal(Dg[k, :o["n"]].view(np.uint32), o["dist"][:o["n"]].view(np.uint32)), k


@pytest.mark.parametrize("margin,frac", [(0.0, 1.0), (0.7, 0.5), (-1.0, 1.0)])
def test_search_learned_equals_explicit_budgets(ctx, idx_cos, margin, frac):
    from puck_amd.engine import search_learned
    ix, q = idx_cos
    gi = _gpu_index(ctx, ix)
    nc, nf, H = 16, 16, 64
    D = 96 + nc + nf
    rng = np.random.default_rng(8)
    pr = _probe(ctx, np.zeros(PM.param_count(D, H), np.float32), nc=nc, nf=nf, hidden=H)
    feats = pr.features(gi, q)
    from puck_amd.probe import init_params
    p = init_params(D, H, feats.mean(0), feats.std(0), 3)
    p[PM.offsets(D, H)["b3"][0]] = 1.5   # budgets around e^1.5
    pr.set_params(p)
    topk = rng.choice([1, 10], q.shape[0]).astype(np.uint32)
    Dl, Il, Rl, B = search_learned(gi, pr, q, 10, 6, 200, margin=margin, min_budget=2, fallback_frac=frac, topk=topk)
    y = pr.predict(feats).astype(np.float64)
    e = np.exp(y + margin)
    want = np.where(e >= frac * 40, 40, np.clip(np.ceil(e), 2, 40)).astype(np.uint32)
    assert np.array_equal(B, want)
    Db, Ib, Rb = gi.search_batch(q, 10, 6, 200, topk=topk, fine_budget=B)
    assert np.array_equal(Rl, Rb) and np.array_equal(Il, Ib) and np.array_equal(Dl.view(np.uint32), Db.view(np.uint32))
    Do, Io, Ro = CI.oracle_index(ix).search_batch_pq(q, 10, 6, 200, topk=topk, fine_budget=B)
    assert np.array_equal(Rl, Ro)
    for k in range(q.shape[0]):
        m = int(Ro[k])
        assert np.array_equal(Il[k, :m], Io[k, :m]) and np.array_equal(Dl[k, :m].view(np.uint32),
                                                                        Do[k, :m].view(np.uint32))


def test_budget_ranks_match_oracle_definition(ctx, idx_cos):
    ix, q = idx_cos
    gi = _gpu_index(ctx, ix)
    o = CI.oracle_index(ix)
    ex = o.export()
    C, F = 40, 40
    st = ex["stationary"].reshape(C, F)
    rng = np.random.default_rng(9)
    cells = rng.integers(0, C * F, (q.shape[0], 5)).astype(np.uint32)
    cells[:, 4] = 0xffffffff
    scc = 6
    r = gi.budget_ranks(q, cells, scc)
    for k in range(q.shape[0]):
        qn = O.normalize(q[k])
        cd = np.array([ex["coarse_norms"][c] - O.dot(ix["coarse"][c], qn) for c in range(C)], np.float32)
        fd = np.array([-O.dot(ix["fine"][f], qn) for f in range(F)], np.float32)
        near = o.search_one(q[k], 10, scc, 100)["coarse"]
        for e in range(5):
            cell = int(cells[k, e])
            if cell == 0xffffffff or cell // F not in set(near.tolist()):
                assert r[k, e] == 0xffffffff
                continue
            c, f = divmod(cell, F)
            keys = _f2ord((cd[c] + st[c]) + fd)
            assert r[k, e] == 1 + int((keys < keys[f]).sum()), (k, e)


def _f2ord(v):
    u = np.ascontiguousarray(v, np.float32).view(np.uint32)
    return np.where(u & 0x80000000, ~u, u | 0x80000000).astype(np.uint32)


def test_budget_labels_refind_plain_hits(ctx, idx_cos):
    import torch
    from puck_amd.probe import budget_labels
    ix, q = idx_cos
    gi = _gpu_index(ctx, ix)
    gt, _ = O.exact_topk(ix["raw"], q, 10)
    qd = torch.from_numpy(q).cuda()
    lab = budget_labels(gi, qd, torch.from_numpy(gt.astype(np.int64)).cuda(),
                        torch.from_numpy(ix["cell_assign"].view(np.int32)).cuda(), 4, 200).cpu().numpy()
    _, Ip, Rp = gi.search(q, 10, 4, 200)
    _, Il, Rl = gi.search_batch(q, 10, 4, 200, fine_budget=lab.astype(np.uint32))
    for k in range(q.shape[0]):
        hits = set(Ip[k, :Rp[k]].tolist()) & set(gt[k, :10].tolist())
        assert hits <= set(Il[k, :Rl[k]].tolist()), k
        assert 1 <= lab[k] <= 40


def test_server_with_probe_matches_search_learned(ctx, idx_cos):
    # graph-captured learned pipeline (budget scan + forked fallback walk) vs
    # the eager fused call
    from puck_amd.engine import Server, search_learned
    from puck_amd.probe import init_params
    ix, q = idx_cos
    gi = _gpu_index(ctx, ix)
    nc, nf, H = 16, 16, 64
    D = 96 + nc + nf
    pr = _probe(ctx, np.zeros(PM.param_count(D, H), np.float32), nc=nc, nf=nf, hidden=H)
    feats = pr.features(gi, q)
    p = init_params(D, H, feats.mean(0), feats.std(0), 5)
    p[PM.offsets(D, H)["b3"][0]] = 2.5
    pr.set_params(p)
    De, Ie, Re, B = search_learned(gi, pr, q, 10, 6, 200, margin=0.0, fallback_frac=0.5)
    assert np.any(B == 40) and np.any(B < 40)   # both the scan and the fallback walk are used
    sv = Server(gi, 10, 6, 200, max_batch=64, probe=pr, margin=0.0, fallback_frac=0.5)
    for s0, bs in ((0, 64), (0, 5), (7, 33)):
        D, I, R = sv.search(q[s0:s0 + bs])
        assert np.array_equal(R, Re[s0:s0 + bs]) and np.array_equal(I, Ie[s0:s0 + bs])
        assert np.array_equal(D.view(np.uint32), De[s0:s0 + bs].view(np.uint32))


def test_probe_save_load_round_trip(ctx, tmp_path):
    from puck_amd import probe as PR
    p = _rand_params(D_, H_, 11)
    pr = _probe(ctx, p)
    path = str(tmp_path / "probe.npz")
    PR.save(pr, path)
    pr2 = PR.load(ctx, path)
    assert (pr2.dim, pr2.nc, pr2.nf, pr2.hidden) == (pr.dim, pr.nc, pr.nf, pr.hidden)
    assert np.array_equal(pr2.get_params().view(np.uint32), pr.get_params().view(np.uint32))
    x = np.random.default_rng(2).standard_normal((37, pr.D)).astype(np.float32)
    assert np.array_equal(pr2.predict(x).view(np.uint32), pr.predict(x).view(np.uint32))
