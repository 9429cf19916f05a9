"""puckgpu_search_noalloc through the C ABI: the search that allocates nothing
(its pass buffers live in the library's per-device scratch array) returns the
bits of puckgpu_search on the same inputs, on two calls in a row (the second
smaller, so it runs over what the first left in the array), and refuses a call
above its limits without touching the outputs."""
import ctypes

import numpy as np
import pytest

from puck_amd import datasets as D
import cpu_index as CI

pytestmark = pytest.mark.gpu


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def test_search_noalloc_matches_search():
    import torch
    from puck_amd import _lib
    from puck_amd.engine import Context, GpuIndex
    x = D.unit_norm_mixture(15000, 96, 300, seed=20, means_seed=21)
    qh = D.unit_norm_mixture(700, 96, 300, seed=22, means_seed=21)
    qh[5] = 0.0  # a zero-norm query: result_num 0 under whether_norm
    ix = CI.build_index(x, 40, 40, 48, whether_norm=True, train_n=6000)
    ctx = Context(0)
    gi = GpuIndex(ctx, ix["coarse"], ix["fine"], ix["codebook"].reshape(-1), ix["cell_assign"], ix["records"],
                  ix["raw"], True)
    L, dev = ctx.L, torch.device("cuda:0")
    q = torch.as_tensor(qh, device=dev)
    for n, scc, ftk, topk in [(700, 20, 1000, 10), (90, 6, 300, 10)]:
        sp = _lib.SearchParams(topk, scc, ftk, 1.0, 0, 0)
        outs = []
        for fn in ("search", "noalloc"):
            D_ = torch.full((n, topk), -1.0, dtype=torch.float32, device=dev)
            I_ = torch.full((n, topk), -1, dtype=torch.int32, device=dev)
            R_ = torch.full((n,), -1, dtype=torch.int32, device=dev)
            torch.cuda.synchronize()
            if fn == "search":
                rc = L.puckgpu_search(ctx.h, gi.h, ctypes.byref(sp), n, _p(q), topk, _p(D_), _p(I_), _p(R_), 1)
            else:
                rc = L.puckgpu_search_noalloc(ctx.h, gi.h, ctypes.byref(sp), n, _p(q), topk, _p(D_), _p(I_), _p(R_))
            assert rc == 0, L.puckgpu_last_error(ctx.h)
            outs.append((D_.cpu().numpy().view(np.uint32), I_.cpu().numpy(), R_.cpu().numpy()))
        (Da, Ia, Ra), (Db, Ib, Rb) = outs
        assert np.array_equal(Ra, Rb), n
        assert Ra[5] == 0
        for k in range(n):
            m = int(Ra[k])
            assert np.array_equal(Ia[k, :m], Ib[k, :m]), (n, k)
            assert np.array_equal(Da[k, :m], Db[k, :m]), (n, k)
    # above the limits: refused, outputs untouched
    sp = _lib.SearchParams(10, 20, 1000, 1.0, 0, 0)
    big = torch.zeros((4097, 96), dtype=torch.float32, device=dev)
    Dz = torch.full((4097, 10), 7.0, dtype=torch.float32, device=dev)
    Iz = torch.full((4097, 10), 7, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    rc = L.puckgpu_search_noalloc(ctx.h, gi.h, ctypes.byref(sp), 4097, _p(big), 10, _p(Dz), _p(Iz), None)
    assert rc == -1 and b"at most 4096" in L.puckgpu_last_error(ctx.h)
    assert bool((Dz == 7.0).all()) and bool((Iz == 7).all())
